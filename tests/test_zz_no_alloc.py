"""The no-allocation exact energy entry point against the allocating one (GPU).

``ddgrs_exact_energy_and_differential_noalloc`` runs the same kernels on the same column
partition with its scratch in a fixed per-device array, so energy, differential and face partials
are bitwise equal to ``ddgrs_exact_energy_and_differential``'s.  Two calls in a row, the second on
a smaller mesh (the array keeps the first call's contents), and a mesh above the size limit is
refused before any work.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def check_equal(pkg, s, exps):
    e0, d0, p0 = pkg.exact_energy_and_differential(s, exps, partials=True)
    e1, d1, p1 = pkg.exact_energy_and_differential_noalloc(s, exps, partials=True)
    print(f"{s.nf} faces: E {e1:.15e} (allocating {e0:.15e}), max |dE diff| {np.abs(d1 - d0).max():.1e}")
    assert e1 == e0
    assert np.array_equal(d1, d0)
    assert np.array_equal(p1, p0)


def test_no_alloc_matches_allocating_entry(pkg, ctx):
    p, f = pkg.generate.config_icosphere_l3()
    s = pkg.Surface(ctx, f, p)
    check_equal(pkg, s, pkg.Exponents())
    # second call, smaller mesh, other exponents (the integer-power kernel)
    p2, f2 = pkg.generate.icosphere(2)
    p2 = pkg.generate.normal_jitter(p2, f2, 0.02, 5)
    s2 = pkg.Surface(ctx, f2, p2)
    check_equal(pkg, s2, pkg.Exponents(4.0, 8.0))
    s2.close()
    s.close()


def test_no_alloc_refuses_large_meshes(pkg, ctx):
    p, f = pkg.generate.icosphere(7)
    s = pkg.Surface(ctx, f, p)
    with pytest.raises(pkg.RsError) as e:
        pkg.exact_energy_and_differential_noalloc(s)
    assert e.value.kind == "InvalidOption" and "131072" in e.value.detail
    s.close()
