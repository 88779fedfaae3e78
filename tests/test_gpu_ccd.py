"""Continuous collision detection on the device (crates/ddg-repulsive-surfaces/src/ccd.rs) against
the FP64 oracle on the same inputs (GPU).

The device walks the reference's box trees in the reference's order and advances every pair with
the same FP64 operations (no contraction), so the comparison is exact: τ_max (bit for bit), the
contact's primitive, the windows and every count (iterations, pairs tested, close, touching and
exempt pairs) — the counts depend on each query's candidate order and horizon, so they pin the
broad phase too.  Intersecting pairs and the band flags: equal arrays.  The step with the CCD cap
runs on the device's own direction (FP32 metric products inside GMRES), so against the oracle's
step it is compared to the solve's tolerance: τ_max within 1e-5 relative and the same contact.
"""
import numpy as np
import pytest

import oracle
from test_oracle_ccd import mean_edge, overlapping_and_apart, soup, sphere, two_triangles, union

pytestmark = pytest.mark.gpu


def device(pkg, ctx, f, p, vel, tau_end, l=1.0, band=None, **kw):
    s = pkg.Surface(ctx, f, p)
    try:
        o = pkg.CcdOptions(reference_length=l, **kw)
        return pkg.first_contact(s, vel, tau_end, o, band=band)
    finally:
        s.close()


def same(dev, ref):
    return (dev.tau_max == ref["tau_max"] and dev.contact == ref["contact"] and dev.gap_abs == ref["gap_abs"]
            and (dev.iterations, dev.pairs_tested, dev.windows, dev.close_pairs, dev.touching_pairs, dev.exempt_pairs)
            == (ref["iterations"], ref["pairs_tested"], ref["windows"], ref["close_pairs"], ref["touching_pairs"],
                ref["exempt_pairs"]))


def check(pkg, ctx, f, p, vel, tau_end, l=1.0, band=None, **kw):
    dev = device(pkg, ctx, f, p, vel, tau_end, l, band, **kw)
    ref = oracle.ccd_first_contact(f, p, vel, tau_end, reference_length=l, band=band, **kw)
    assert same(dev, ref), (dev, ref)
    return dev


def test_reference_scenes(pkg, ctx):
    scenes = [
        (two_triangles([[0.2, 0.2, 1.0], [1.0, 1.0, 1.5], [-0.5, 1.2, 1.5]], [0.0, 0.0, -1.0]), 2.0),
        (two_triangles([[0.5, -0.5, 1.0], [0.5, 0.5, 1.0], [3.0, 0.0, 1.0]], [0.0, 0.0, -1.0]), 2.0),
        (two_triangles([[0.2, 0.2, 1.0], [1.0, 1.0, 1.0], [-0.5, 1.2, 1.0]], [3.0, 0.0, 0.0]), 5.0),
        (two_triangles([[2.0, 0.5, 0.0], [3.0, 0.5, 0.0], [2.5, 1.5, 0.0]], [-4.0, -1.5, 0.0]), 1.0),
        (two_triangles([[0.2, 0.2, 0.02], [1.0, 1.0, 0.5], [-0.5, 1.2, 0.5]], [0.0, 0.0, -1.0]), 2.0),
        (two_triangles([[0.2, 0.2, 0.02], [1.0, 1.0, 0.5], [-0.5, 1.2, 0.5]], [0.0, 0.0, 1.0]), 2.0),
        (two_triangles([[0.2, 0.2, 1.0], [1.0, 1.0, 1.0], [-0.5, 1.2, 1.0]], [0.0, 0.0, 0.0]), 3.0),
    ]
    for (f, p, vel), t in scenes:
        check(pkg, ctx, f, p, vel, t)
    f, p, vel = two_triangles([[0.2, 0.2, 0.0], [1.0, 1.0, 0.5], [-0.5, 1.2, 0.5]], [0.0, 0.0, 1.0])
    r = check(pkg, ctx, f, p, vel, 1.0)
    assert r.touching_pairs == 1 and r.tau_max == 1.0
    f, p = soup([[[-1.0, 0.0, 0.0], [1.0, 0.0, 0.0], [0.0, -0.01, 0.0]],
                 [[0.005, -1.0, 1.0], [0.005, 1.0, 1.0], [0.015, 0.2, 1.0]]])
    vel = np.zeros((6, 3))
    vel[3:] = [0.0, 0.0, -1.0]
    r = check(pkg, ctx, f, p, vel, 2.0)
    assert r.contact[1] == "edge_edge" and 0.94 < r.tau_max <= 0.95 + 1e-9


def test_random_soups(pkg, ctx):
    rng = np.random.default_rng(2024)
    contacts = 0
    for _ in range(24):
        k = 8
        c = rng.uniform(0.0, 2.0, (k, 1, 3))
        f, p = soup(c + rng.uniform(-0.4, 0.4, (k, 3, 3)))
        vel = rng.uniform(-1.0, 1.0, (3 * k, 3))
        r = check(pkg, ctx, f, p, vel, 0.6, mean_edge(f, p))
        contacts += r.contact is not None
    assert contacts > 0


def test_sphere_pairs_and_leaf_sizes(pkg, ctx):
    rng = np.random.default_rng(77)
    for leaf in (8, 1, 5):
        a = sphere(2, [0.0, 0.0, 0.0], 1.0)
        b = sphere(2, [rng.uniform(-0.5, 0.5), rng.uniform(-0.5, 0.5), 2.6], rng.uniform(0.7, 1.0))
        f, p = union(a, b)
        vel = np.zeros_like(p)
        vel[len(a[1]):] = [0.3, -0.2, -1.3]
        r = check(pkg, ctx, f, p, vel, 2.5, mean_edge(f, p), leaf_size=leaf)
        assert r.contact is not None and r.windows >= 1


def test_large_surface_smooth_motion(pkg, ctx):
    """A level-4 icosphere (5,120 faces) under a smooth velocity field that folds it: every count
    over some twenty thousand queries equals the oracle's."""
    p, f = pkg.generate.icosphere(4)
    p = np.asarray(p)
    f = np.asarray(f, dtype=np.uint32)
    vel = np.stack([np.sin(3 * p[:, 1]), np.cos(2 * p[:, 2]), -2.0 * p[:, 2] * np.abs(p[:, 0])], axis=1)
    r = check(pkg, ctx, f, p, vel, 5.0, None)
    print(f"L4: tau_max {r.tau_max:.6g} contact {r.contact} pairs {r.pairs_tested} iterations {r.iterations} "
          f"windows {r.windows}")
    assert r.contact is not None


def test_intersecting_pairs_and_band(pkg, ctx):
    (f, p), n = overlapping_and_apart()
    s = pkg.Surface(ctx, f, p)
    pairs = pkg.intersecting_pairs(s)
    ref = oracle.intersecting_pairs(f, p)
    assert len(ref) > 0 and np.array_equal(pairs, ref)
    flags, npairs, nb = pkg.intersection_band(s)
    rflags, rnp, rnb = oracle.intersection_band(f, p)
    assert np.array_equal(flags, rflags) and (npairs, nb) == (rnp, rnb)
    assert np.array_equal(s.edges(), np.array(sorted({(min(a, b), max(a, b)) for t in f.tolist()
                                                      for a, b in ((t[0], t[1]), (t[1], t[2]), (t[2], t[0]))}),
                                               dtype=np.uint32))
    s.close()
    # coplanar overlap found; an embedded pair of spheres has none
    fc, pc = soup([[[0.0, 0.0, 0.0], [1.0, 0.0, 0.0], [0.0, 1.0, 0.0]], [[0.2, 0.2, 0.0], [1.2, 0.2, 0.0], [0.2, 1.2, 0.0]]])
    s = pkg.Surface(ctx, fc, pc)
    assert pkg.intersecting_pairs(s).tolist() == [[0, 1]]
    s.close()
    m = union(sphere(2, [0, 0, 0], 1.0), sphere(2, [4.0, 0, 0], 1.0))
    s = pkg.Surface(ctx, *m)
    assert len(pkg.intersecting_pairs(s)) == 0 and pkg.intersection_band(s)[1:] == (0, 0)
    s.close()


def test_band_first_contact(pkg, ctx):
    (f, p), n = overlapping_and_apart()
    l = mean_edge(f, p)
    band, _, _ = oracle.intersection_band(f, p)
    vel = np.zeros_like(p)
    vel[n:2 * n] = [0.0, 0.0, 0.4]
    banded = check(pkg, ctx, f, p, vel, 1.0, l, band=band)
    plain = check(pkg, ctx, f, p, vel, 1.0, l)
    assert banded.exempt_pairs > 0 and banded.tau_max > plain.tau_max
    vel = np.zeros_like(p)
    vel[2 * n:3 * n] = [0.0, 0.0, 1.0]
    r = check(pkg, ctx, f, p, vel, 3.0, l, band=band)
    assert r.contact is not None


def test_option_and_input_errors(pkg, ctx):
    f, p, vel = two_triangles([[0.2, 0.2, 1.0], [1.0, 1.0, 1.0], [-0.5, 1.2, 1.0]], [0.0, 0.0, 0.0])
    s = pkg.Surface(ctx, f, p)
    for kw in [dict(gap=0.0), dict(gap=float("nan")), dict(max_iterations=0), dict(close_factor=1.0),
               dict(leaf_size=0), dict(reference_length=-1.0)]:
        with pytest.raises(pkg.RsError):
            pkg.first_contact(s, vel, 1.0, pkg.CcdOptions(**kw))
    with pytest.raises(pkg.RsError):
        pkg.first_contact(s, vel[:5], 1.0)
    with pytest.raises(pkg.RsError):
        pkg.first_contact(s, vel, -1.0)
    v = vel.copy()
    v[2, 1] = np.nan
    with pytest.raises(pkg.RsError):
        pkg.first_contact(s, v, 1.0)
    bad = p.copy()
    bad[4] = bad[3]
    with pytest.raises(pkg.RsError):
        pkg.first_contact(s, vel, 1.0, start=bad)
    s.close()


def test_step_with_ccd_cap(pkg, ctx):
    """The torus's flow direction folds the inner ring: with a large absolute start the CCD caps
    τ₀ at ⅔ τ_max (session.rs:950-952).  Device step against the oracle's step (same options)."""
    p, f = pkg.generate.torus(24, 12, 1.0, 0.45)
    p = np.asarray(p)
    f = np.asarray(f, dtype=np.uint32)
    ref = oracle.Session(f, p, start=("abs", 1e3), ccd={}).step()
    s = pkg.Surface(ctx, f, p)
    ses = pkg.Session(s, pkg.StepOptions(start=pkg.StartRule.absolute(1e3), ccd=pkg.CcdOptions()))
    r = ses.step()
    print(f"torus: device tau_ccd {r.tau_ccd} contact {r.ccd.contact}; oracle {ref['ccd']['tau_max']} "
          f"{ref['ccd']['contact']}")
    assert r.ccd is not None and r.ccd.contact is not None
    assert r.ccd.contact[1:] == ref["ccd"]["contact"][1:]
    assert abs(r.tau_ccd - ref["ccd"]["tau_max"]) <= 1e-5 * ref["ccd"]["tau_max"]
    assert r.tau0 == pytest.approx(2.0 / 3.0 * r.tau_ccd, rel=1e-15)
    assert abs(r.tau - ref["tau"]) <= 1e-5 * ref["tau"]
    assert r.accepted
    # without the CCD the first trial is the absolute start
    s.set_positions(p)
    ses2 = pkg.Session(s, pkg.StepOptions(start=pkg.StartRule.absolute(1e3)))
    r2 = ses2.step()
    assert r2.tau0 == 1e3 and r2.ccd is None and r2.tau_ccd is None
    ses.close()
    ses2.close()
    s.close()
