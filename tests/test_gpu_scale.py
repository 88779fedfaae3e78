"""The level-7 scale configuration (163,842 vertices, 327,680 faces, 64 bumps) on the device (GPU).

Three consecutive plain Hs steps with Barnes-Hut θ = 0.25: every step's GMRES converges to the
true relative residual 1e-6, the energy decreases, and the energy at each step's start (the
device's Barnes-Hut energy on lists built at those positions) equals the FP64 oracle's Barnes-Hut
energy at the same positions to 1e-5 relative.  The 20-step run with the drift after steps 1, 10
and 20 is `bench.py --config l7`.
"""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu


def test_l7_steps(pkg, ctx):
    p, f = pkg.generate.config_bumpy_l7()
    s = pkg.Surface(ctx, f, p)
    ses = pkg.Session(s, pkg.StepOptions(theta=0.25))
    prev = None
    for k in range(3):
        x = s.positions()
        r = ses.step()
        e0 = oracle.BhTree(f, x, theta=0.25, threads=16).energy()
        rel = (r.energy_before - e0) / e0
        print(f"L7 step {k}: E {r.energy_before:.12g} / oracle {e0:.12g} (rel {rel:.2e}), gmres {r.gmres.iterations}, "
              f"seconds {r.seconds}")
        assert abs(rel) <= 1e-5
        assert r.gmres.converged and r.gmres.relative_residual <= 1e-6
        assert r.energy_trial < r.energy_before
        if prev is not None:
            assert r.energy_before < prev
        prev = r.energy_before
