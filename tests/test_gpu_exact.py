"""Parity of the HIP all-pairs energy kernel with the FP64 oracle (GPU).

Tolerances (FP32 pair kernel with compensated accumulation against the FP64 oracle):
energy ≤ 1e-5 relative, differential ≤ 1e-4 relative in max norm.
"""
import numpy as np
import pytest

import oracle
from conftest import rel_max
from golden import fixtures as fx

pytestmark = pytest.mark.gpu

E_TOL = 1e-5
G_TOL = 1e-4


@pytest.mark.parametrize("name", [n for n in fx.cases("energy") if n != "rs-flat-energy"])
def test_lean_energy_cases(pkg, ctx, name):
    c = fx.read_case(name)
    a, b = c["exponents"]
    s = pkg.Surface(ctx, c["faces"], c["positions"])
    e, d = pkg.exact_energy_and_differential(s, pkg.Exponents(a, b))
    e0, d0, _ = oracle.exact(c["faces"], c["positions"], a, b)
    print(f"{name}: E rel {abs(e - e0) / e0:.2e}, dE max-rel {rel_max(d, d0):.2e}")
    assert abs(e - e0) <= E_TOL * e0
    assert rel_max(d, d0) <= G_TOL


def test_flat_patch_is_exactly_zero(pkg, ctx):
    c = fx.read_case("rs-flat-energy")
    s = pkg.Surface(ctx, c["faces"], c["positions"])
    e, d = pkg.exact_energy_and_differential(s)
    assert e == 0.0 and np.all(d == 0.0)


def test_icosphere_l3_config(pkg, ctx):
    p, f = pkg.generate.config_icosphere_l3()
    s = pkg.Surface(ctx, f, p)
    e, d, fp = pkg.exact_energy_and_differential(s, partials=True)
    e0, d0, fp0 = oracle.exact(f, p)
    print(f"L3: E {e:.15e} oracle {e0:.15e} rel {abs(e - e0) / e0:.2e}; dE max-rel {rel_max(d, d0):.2e}; "
          f"partials max-rel {rel_max(fp, fp0):.2e}")
    assert abs(e - e0) <= E_TOL * e0
    assert rel_max(d, d0) <= G_TOL
    e2, d2 = pkg.exact_energy_and_differential(s)
    assert e2 == e and np.array_equal(d2, d), "run-to-run reproducible"


def test_torus_16k(pkg, ctx):
    p, f = pkg.generate.config_torus_16k()
    s = pkg.Surface(ctx, f, p)
    e, d = pkg.exact_energy_and_differential(s)
    e0, d0, _ = oracle.exact(f, p)
    print(f"T16k: E rel {abs(e - e0) / e0:.2e}; dE max-rel {rel_max(d, d0):.2e}")
    assert abs(e - e0) <= E_TOL * e0
    assert rel_max(d, d0) <= G_TOL


def test_bumpy_l6_identities(pkg, ctx):
    """At 81,920 faces (too large for the oracle in a test): scaling, Σg = 0 and the Euler
    identity Σ f·g = (4 + α − β) E, to FP32 accuracy."""
    p, f = pkg.generate.config_bumpy_l6()
    s = pkg.Surface(ctx, f, p)
    e, d = pkg.exact_energy_and_differential(s)
    s.set_positions(1.5 * p)
    e2, _ = pkg.exact_energy_and_differential(s)
    assert abs(e2 - 1.5 ** -2 * e) <= 1e-5 * e
    gs = np.abs(d).max()
    assert np.abs(d.sum(axis=0)).max() <= 1e-5 * gs * np.sqrt(len(p))
    assert abs((p * d).sum() + 2.0 * e) <= 1e-5 * e


@pytest.mark.parametrize(
    "name,kind",
    [("rs-error-same-barycenter", "NonFiniteEnergy"), ("rs-error-degenerate", "DegenerateFace"),
     ("rs-error-exponents", "InvalidExponents")],
)
def test_errors(pkg, ctx, name, kind):
    c = fx.read_case(name)
    a, b = c["exponents"]
    with pytest.raises(pkg.RsError) as ei:
        s = pkg.Surface(ctx, c["faces"], c["positions"])
        pkg.exact_energy_and_differential(s, pkg.Exponents(a, b))
    assert ei.value.kind == kind
