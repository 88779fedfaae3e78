"""Barnes–Hut on the device against the FP64 oracle (GPU).

* the face BVH (order, node ranges, children) and the build-time cluster data are bitwise the
  oracle's, and so are the interaction lists (same θ, same leaf size);
* Ẽ within 1e-5 relative and ∂Ẽ within 1e-4 relative in max norm of the oracle's Barnes–Hut
  evaluation on the same tree and lists (FP32 pair kernels, compensated / FP64 epilogues);
* θ = 0 reproduces the exact all-pairs energy; refit follows the oracle's refit.
"""
import numpy as np
import pytest

import oracle
from conftest import rel_max

pytestmark = pytest.mark.gpu


def compare_structure(bh, ob):
    t, to = bh.tree(), ob.tree()
    for k in ("order", "start", "end", "right"):
        assert np.array_equal(t[k], to[k]), k
    # build-time cluster data: area, centroid, projector, box, radius — bitwise
    assert np.array_equal(t["clusters"], to["clusters"])
    L, Lo = bh.lists(), ob.lists()
    for k in ("lumped_off", "lumped", "exact_off", "exact"):
        assert np.array_equal(L[k].astype(np.int64), Lo[k].astype(np.int64)), k
    assert bh.counts() == ob.counts()


@pytest.mark.parametrize("theta", [0.25, 0.5])
def test_torus_16k(pkg, ctx, theta):
    p, f = pkg.generate.config_torus_16k()
    s = pkg.Surface(ctx, f, p)
    bh = pkg.BhTree(s, pkg.BhOptions(theta, 8))
    ob = oracle.BhTree(f, p, theta=theta, leaf_size=8)
    compare_structure(bh, ob)
    e, d = bh.energy_and_differential()
    e0, d0, _ = ob.energy_and_differential()
    print(f"T16k theta {theta}: E rel {abs(e - e0) / e0:.2e}, dE max-rel {rel_max(d, d0):.2e}, counts {bh.counts()}")
    assert abs(e - e0) <= 1e-5 * e0
    assert rel_max(d, d0) <= 1e-4
    ev = bh.energy()
    assert abs(ev - e) <= 1e-6 * e
    ex, _, _ = oracle.exact(f, p)
    print(f"  vs FP64 all pairs: E rel {abs(e - ex) / ex:.2e}")


def test_theta_zero_equals_exact(pkg, ctx):
    p, f = pkg.generate.config_icosphere_l3()
    s = pkg.Surface(ctx, f, p)
    bh = pkg.BhTree(s, pkg.BhOptions(0.0, 8))
    e, d = bh.energy_and_differential()
    e0, d0, _ = oracle.exact(f, p)
    assert abs(e - e0) <= 1e-5 * e0
    assert rel_max(d, d0) <= 1e-4


def test_refit_follows_the_oracle(pkg, ctx):
    p, f = pkg.generate.config_torus_16k()
    s = pkg.Surface(ctx, f, p)
    bh = pkg.BhTree(s, pkg.BhOptions(0.25, 8))
    ob = oracle.BhTree(f, p, theta=0.25)
    q = p + 0.002 * np.sin(7.0 * p[:, [1, 2, 0]])
    bh.refit(q)
    ob.refit(q)
    e, d = bh.energy_and_differential()
    e0, d0, _ = ob.energy_and_differential()
    assert abs(e - e0) <= 1e-5 * e0
    assert rel_max(d, d0) <= 1e-4
    assert abs(bh.energy() - ob.energy()) <= 1e-5 * e0


def test_bumpy_l6(pkg, ctx):
    p, f = pkg.generate.config_bumpy_l6()
    s = pkg.Surface(ctx, f, p)
    bh = pkg.BhTree(s, pkg.BhOptions(0.5, 8))
    ob = oracle.BhTree(f, p, theta=0.5, leaf_size=8)
    compare_structure(bh, ob)
    e, d = bh.energy_and_differential()
    e0, d0, _ = ob.energy_and_differential()
    print(f"L6 theta 0.5: E rel {abs(e - e0) / e0:.2e}, dE max-rel {rel_max(d, d0):.2e}")
    assert abs(e - e0) <= 1e-5 * e0
    assert rel_max(d, d0) <= 1e-4


def test_device_lists_equal_host_lists(pkg, ctx, monkeypatch):
    """Lists and transposes built on the device give bitwise the evaluation of the host-built ones."""
    p, f = pkg.generate.config_torus_16k()
    s = pkg.Surface(ctx, f, p)
    dev = pkg.BhTree(s, pkg.BhOptions(0.5, 8))
    monkeypatch.setenv("DDGRS_BH_LISTS_ON_HOST", "1")
    host = pkg.BhTree(s, pkg.BhOptions(0.5, 8))
    monkeypatch.delenv("DDGRS_BH_LISTS_ON_HOST")
    assert dev.counts() == host.counts()
    Ld, Lh = dev.lists(), host.lists()
    for k in Ld:
        assert np.array_equal(Ld[k], Lh[k]), k
    e, d = dev.energy_and_differential()
    eh, dh = host.energy_and_differential()
    assert e == eh
    assert np.array_equal(d, dh)
