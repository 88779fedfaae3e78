"""The oracle's continuous collision detection (crates/ddg-repulsive-surfaces/src/ccd.rs) against
the reference's own CCD tests (tests/ccd.rs), closed forms of the curves crate's distances and
brute force.

* Distances (ddg-repulsive-curves/src/ccd.rs:236-320): closed forms (height above a face, the
  separation of skew segments, endpoint cases) and a dense sampling of both primitives, which the
  exact distance must not exceed and must approach.
* first_contact (ccd.rs:781): the scenes of tests/ccd.rs — head-on vertex-face and edge-edge
  contacts at t = 0.95 (gap 0.05 at unit speed), sliding and stationary scenes without contact,
  initially close pairs (closing to half their distance) and touching pairs, option errors —
  and random soups / sphere pairs whose sampled distances stay above the gap on [0, τ_max].
* Static queries (ccd.rs:1153-1420): min_distance equals brute force; penetrating spheres
  intersect; coplanar overlaps are found and coplanar neighbours are not; the band of two
  overlapping spheres holds only their vertices and frees their sliding while a new contact of
  two other spheres is still found as for those spheres alone.
"""
import numpy as np
import pytest

import oracle
from conftest import load_package


def soup(tris):
    tris = np.asarray(tris, dtype=np.float64)
    return np.arange(3 * len(tris), dtype=np.uint32).reshape(-1, 3), tris.reshape(-1, 3).copy()


def union(a, b):
    fa, pa = a
    fb, pb = b
    return np.vstack([fa, fb + len(pa)]).astype(np.uint32), np.vstack([pa, pb])


def sphere(level, center, radius):
    p, f = load_package().generate.icosphere(level)
    return np.asarray(f, dtype=np.uint32), np.asarray(center) + radius * np.asarray(p)


def two_triangles(b, vb):
    f, p = soup([[[0.0, 0.0, 0.0], [1.0, 0.0, 0.0], [0.0, 1.0, 0.0]], b])
    vel = np.zeros((6, 3))
    vel[3:] = vb
    return f, p, vel


def mean_edge(f, p):
    e = {(min(a, b), max(a, b)) for t in f.tolist() for a, b in ((t[0], t[1]), (t[1], t[2]), (t[2], t[0]))}
    return float(np.mean([np.linalg.norm(p[a] - p[b]) for a, b in e]))


def fc(f, p, vel, tau_end, l=1.0, **kw):
    return oracle.ccd_first_contact(f, p, vel, tau_end, reference_length=l, **kw)


# ---- distances ---------------------------------------------------------------------------


def test_distance_closed_forms():
    d = oracle.ccd_primitive
    a, b, c = [0.0, 0.0, 0.0], [1.0, 0.0, 0.0], [0.0, 1.0, 0.0]
    assert d("point_triangle", [0.2, 0.3, 0.7], a, b, c) == pytest.approx(0.7, abs=1e-15)
    assert d("point_triangle", [0.2, 0.3, -0.4], a, b, c) == pytest.approx(0.4, abs=1e-15)
    # beyond the hypotenuse: the distance to that edge
    assert d("point_triangle", [1.0, 1.0, 0.0], a, b, c) == pytest.approx(np.sqrt(0.5), abs=1e-15)
    # beyond vertex b along x: the distance to b
    assert d("point_triangle", [3.0, -1.0, 2.0], a, b, c) == pytest.approx(3.0, abs=1e-15)
    # skew segments crossing at a right angle, 0.5 apart
    assert d("segment_segment", [-1, 0, 0], [1, 0, 0], [0, -1, 0.5], [0, 1, 0.5]) == pytest.approx(0.5, abs=1e-15)
    # parallel segments offset by (0, 2, 0) and overlapping in x
    assert d("segment_segment", [0, 0, 0], [2, 0, 0], [1, 2, 0], [3, 2, 0]) == pytest.approx(2.0, abs=1e-15)
    # collinear, disjoint: the gap between the endpoints
    assert d("segment_segment", [0, 0, 0], [1, 0, 0], [1.5, 0, 0], [3, 0, 0]) == pytest.approx(0.5, abs=1e-15)
    assert d("point_segment", [0.5, 1.0, 0.0], a, b) == pytest.approx(1.0, abs=1e-15)
    assert d("point_segment", [-1.0, 0.0, 0.0], a, b) == pytest.approx(1.0, abs=1e-15)


def test_distances_against_sampling():
    rng = np.random.default_rng(5)
    s = np.linspace(0.0, 1.0, 201)
    uu, vv = np.meshgrid(s, s)
    m = uu + vv <= 1.0
    bu, bv = uu[m], vv[m]
    for _ in range(40):
        p, a, b, c = rng.standard_normal((4, 3))
        pts = a + bu[:, None] * (b - a) + bv[:, None] * (c - a)
        sampled = np.min(np.linalg.norm(pts - p, axis=1))
        exact = oracle.ccd_primitive("point_triangle", p, a, b, c)
        assert exact <= sampled + 1e-12
        assert exact >= sampled - 0.02 * np.linalg.norm(b - a) - 0.02 * np.linalg.norm(c - a)
        p1, q1, p2, q2 = rng.standard_normal((4, 3))
        x = p1 + s[:, None] * (q1 - p1)
        y = p2 + s[:, None] * (q2 - p2)
        sampled = np.min(np.linalg.norm(x[:, None, :] - y[None, :, :], axis=2))
        exact = oracle.ccd_primitive("segment_segment", p1, q1, p2, q2)
        assert exact <= sampled + 1e-12
        assert exact >= sampled - 0.01 * (np.linalg.norm(q1 - p1) + np.linalg.norm(q2 - p2))


# ---- first contact: the reference's scenes -----------------------------------------------


def test_head_on_vertex_face():
    f, p, vel = two_triangles([[0.2, 0.2, 1.0], [1.0, 1.0, 1.5], [-0.5, 1.2, 1.5]], [0.0, 0.0, -1.0])
    r = fc(f, p, vel, 2.0)
    assert 0.94 < r["tau_max"] <= 0.95 + 1e-9
    assert r["contact"][1:] == ("vertex_triangle", 3, 0)
    assert (r["close_pairs"], r["touching_pairs"]) == (0, 0)


def test_edge_edge_crossing():
    f, p, vel = two_triangles([[0.5, -0.5, 1.0], [0.5, 0.5, 1.0], [3.0, 0.0, 1.0]], [0.0, 0.0, -1.0])
    r = fc(f, p, vel, 2.0)
    assert 0.94 < r["tau_max"] <= 0.95 + 1e-9
    f, p = soup([[[-1.0, 0.0, 0.0], [1.0, 0.0, 0.0], [0.0, -0.01, 0.0]],
                 [[0.005, -1.0, 1.0], [0.005, 1.0, 1.0], [0.015, 0.2, 1.0]]])
    vel = np.zeros((6, 3))
    vel[3:] = [0.0, 0.0, -1.0]
    r = fc(f, p, vel, 2.0)
    assert 0.94 < r["tau_max"] <= 0.95 + 1e-9
    assert r["contact"][1] == "edge_edge"


def test_sliding_and_stationary():
    f, p, vel = two_triangles([[0.2, 0.2, 1.0], [1.0, 1.0, 1.0], [-0.5, 1.2, 1.0]], [3.0, 0.0, 0.0])
    r = fc(f, p, vel, 5.0)
    assert r["tau_max"] == 5.0 and r["contact"] is None
    f, p, vel = two_triangles([[2.0, 0.5, 0.0], [3.0, 0.5, 0.0], [2.5, 1.5, 0.0]], [-4.0, -1.5, 0.0])
    r = fc(f, p, vel, 1.0)
    assert r["tau_max"] < 1.0 or r["contact"] is None
    f, p, _ = two_triangles([[0.2, 0.2, 1.0], [1.0, 1.0, 1.0], [-0.5, 1.2, 1.0]], [0.0, 0.0, 0.0])
    r = fc(f, p, np.zeros((6, 3)), 3.0)
    assert r["tau_max"] == 3.0 and r["contact"] is None and r["windows"] == 1
    r = fc(f, p, np.tile([1.0, 2.0, 3.0], (6, 1)), 3.0)
    assert r["tau_max"] == 3.0


def test_initially_close_and_touching_pairs():
    f, p, vel = two_triangles([[0.2, 0.2, 0.02], [1.0, 1.0, 0.5], [-0.5, 1.2, 0.5]], [0.0, 0.0, -1.0])
    r = fc(f, p, vel, 2.0)
    assert r["close_pairs"] == 1
    assert 0.009 < r["tau_max"] <= 0.01 + 1e-9
    r = fc(f, p, -vel, 2.0)
    assert r["tau_max"] == 2.0 and r["close_pairs"] == 1
    f, p, vel = two_triangles([[0.2, 0.2, 0.0], [1.0, 1.0, 0.5], [-0.5, 1.2, 0.5]], [0.0, 0.0, 1.0])
    r = fc(f, p, vel, 1.0)
    assert r["touching_pairs"] == 1 and r["tau_max"] == 1.0


def test_option_and_input_errors():
    f, p, vel = two_triangles([[0.2, 0.2, 1.0], [1.0, 1.0, 1.0], [-0.5, 1.2, 1.0]], [0.0, 0.0, 0.0])
    for kw in [dict(gap=0.0), dict(gap=np.nan), dict(max_iterations=0), dict(close_factor=1.0), dict(leaf_size=0),
               dict(reference_length=-1.0)]:
        with pytest.raises(oracle.OracleError):
            oracle.ccd_first_contact(f, p, vel, 1.0, **kw)
    with pytest.raises(oracle.OracleError):
        fc(f, p, vel[:5], 1.0)
    with pytest.raises(oracle.OracleError):
        fc(f, p, vel, -1.0)
    v = vel.copy()
    v[2, 1] = np.nan
    with pytest.raises(oracle.OracleError):
        fc(f, p, v, 1.0)


# ---- sampled scenes ------------------------------------------------------------------------


def sampled(f, p, vel, t_end, n, level):
    """(first sampled time below `level`, smallest sampled distance) over n + 1 samples."""
    first, dmin = None, np.inf
    for i in range(n + 1):
        t = t_end * i / n
        x = p + t * vel
        d = 0.0 if len(oracle.intersecting_pairs(f, x)) else oracle.min_distance(f, x)
        dmin = min(dmin, d)
        if first is None and d < level:
            first = t
    return first, dmin


def check_scene(f, p, vel, t_end, l, samples):
    gap = 0.05 * l
    r = fc(f, p, vel, t_end, l)
    assert 0.0 <= r["tau_max"] <= t_end
    _, dmin = sampled(f, p, vel, r["tau_max"], samples, 0.0)
    assert dmin >= gap * (1.0 - 1e-6), (r, dmin)
    below, _ = sampled(f, p, vel, t_end, samples, 1.2 * gap)
    if below is None:
        assert r["tau_max"] == t_end
    contact, _ = sampled(f, p, vel, t_end, samples, 1e-300)
    if contact is not None:
        assert r["tau_max"] <= contact + 1e-12
    return r["tau_max"], contact


def random_soup(rng, k):
    while True:
        c = rng.uniform(0.0, 2.0, (k, 1, 3))
        tris = c + rng.uniform(-0.4, 0.4, (k, 3, 3))
        f, p = soup(tris)
        vel = rng.uniform(-1.0, 1.0, (3 * k, 3))
        l = mean_edge(f, p)
        if not len(oracle.intersecting_pairs(f, p)) and oracle.min_distance(f, p) > 1.5 * 0.05 * l:
            return f, p, vel, l


def test_random_soups_against_sampling():
    rng = np.random.default_rng(1000)
    results = []
    for _ in range(60):
        f, p, vel, l = random_soup(rng, 8)
        results.append(check_scene(f, p, vel, 0.6, l, 150))
    with_contact = [(t, c) for t, c in results if c is not None and c > 0.0]
    assert len(with_contact) >= 5
    ratios = sorted(t / c for t, c in with_contact)
    assert ratios[len(ratios) // 2] > 0.5


def test_sphere_pairs_against_sampling():
    rng = np.random.default_rng(77)
    any_contact = False
    for _ in range(6):
        a = sphere(1, [0.0, 0.0, 0.0], 1.0)
        c = [rng.uniform(-0.5, 0.5), rng.uniform(-0.5, 0.5), 2.6]
        b = sphere(1, c, rng.uniform(0.7, 1.0))
        f, p = union(a, b)
        speed = rng.uniform(0.5, 2.0)
        tilt = rng.uniform(-0.3, 0.3, 2)
        vel = np.zeros_like(p)
        vel[len(a[1]):] = [speed * tilt[0], speed * tilt[1], -speed]
        l = mean_edge(f, p)
        _, contact = check_scene(f, p, vel, 2.5, l, 60)
        any_contact = any_contact or contact is not None
    assert any_contact


# ---- static queries and the band ---------------------------------------------------------


def brute_min(f, p):
    best = np.inf
    for i in range(len(f)):
        for j in range(len(f)):
            for v in f[i]:
                if v in f[j]:
                    continue
                best = min(best, oracle.ccd_primitive("point_triangle", p[v], *p[f[j]]))
    e = sorted({(min(a, b), max(a, b)) for t in f for a, b in ((t[0], t[1]), (t[1], t[2]), (t[2], t[0]))})
    for i in range(len(e)):
        for j in range(i + 1, len(e)):
            if set(e[i]) & set(e[j]):
                continue
            best = min(best, oracle.ccd_primitive("segment_segment", p[e[i][0]], p[e[i][1]], p[e[j][0]], p[e[j][1]]))
    return best


def test_static_checks():
    a = sphere(1, [0, 0, 0], 1.0)
    m = union(a, sphere(1, [0, 0, 4.0], 1.0))
    assert len(oracle.intersecting_pairs(*m)) == 0
    assert abs(oracle.min_distance(*m) - brute_min(*m)) < 1e-12
    m2 = union(a, sphere(1, [0, 0, 1.5], 1.0))
    pairs = oracle.intersecting_pairs(*m2)
    assert len(pairs) and all(tuple(pairs[i]) < tuple(pairs[i + 1]) for i in range(len(pairs) - 1))
    assert len(oracle.intersecting_pairs(*a)) == 0 and oracle.min_distance(*a) > 0.0
    rng = np.random.default_rng(9)
    f, p = soup(rng.uniform(0.0, 2.0, (6, 1, 3)) + rng.uniform(-0.4, 0.4, (6, 3, 3)))
    assert abs(oracle.min_distance(f, p) - brute_min(f, p)) < 1e-12


def test_coplanar_faces():
    f, p = soup([[[0.0, 0.0, 0.0], [1.0, 0.0, 0.0], [0.0, 1.0, 0.0]],
                 [[0.2, 0.2, 0.0], [1.2, 0.2, 0.0], [0.2, 1.2, 0.0]]])
    assert oracle.intersecting_pairs(f, p).tolist() == [[0, 1]]
    # a flat grid in generic planes: neighbours with one common vertex never intersect
    n = 12
    idx = np.arange((n + 1) ** 2).reshape(n + 1, n + 1)
    faces = []
    for i in range(n):
        for j in range(n):
            faces += [[idx[i, j], idx[i + 1, j], idx[i + 1, j + 1]], [idx[i, j], idx[i + 1, j + 1], idx[i, j + 1]]]
    faces = np.array(faces, dtype=np.uint32)
    g = np.stack(np.meshgrid(np.linspace(0, 1, n + 1), np.linspace(0, 1, n + 1), indexing="ij"), -1).reshape(-1, 2)
    rng = np.random.default_rng(7)
    for _ in range(20):
        u, w = rng.uniform(-0.5, 0.5, (2, 3))
        x = 0.3 + g[:, :1] * u + g[:, 1:] * w
        assert len(oracle.intersecting_pairs(faces, x)) == 0


def overlapping_and_apart():
    a = sphere(2, [0, 0, 0], 1.0)
    b = sphere(2, [1.2, 0, 0], 1.0)
    c = sphere(2, [0, 0, 6.0], 1.0)
    d = sphere(2, [0, 0, 9.0], 1.0)
    n = len(a[1])
    return union(union(union(a, b), c), d), n


def test_band():
    m = union(sphere(2, [0, 0, 0], 1.0), sphere(2, [4.0, 0, 0], 1.0))
    flags, npairs, nb = oracle.intersection_band(*m)
    assert npairs == 0 and nb == 0 and not flags.any()
    (f, p), n = overlapping_and_apart()
    flags, npairs, nb = oracle.intersection_band(f, p)
    assert npairs > 0 and nb == int(flags.sum()) and nb > 0
    assert not flags[2 * n:].any()
    l = mean_edge(f, p)
    vel = np.zeros_like(p)
    vel[n:2 * n] = [0.0, 0.0, 0.4]
    plain = fc(f, p, vel, 1.0, l)
    banded = fc(f, p, vel, 1.0, l, band=flags)
    assert banded["exempt_pairs"] > 0 and banded["tau_max"] > plain["tau_max"]
    vel = np.zeros_like(p)
    vel[2 * n:3 * n] = [0.0, 0.0, 1.0]
    banded = fc(f, p, vel, 3.0, l, band=flags)
    fcd, pcd = union(sphere(2, [0, 0, 6.0], 1.0), sphere(2, [0, 0, 9.0], 1.0))
    vcd = np.zeros_like(pcd)
    vcd[:n] = [0.0, 0.0, 1.0]
    alone = fc(fcd, pcd, vcd, 3.0, l)
    assert alone["contact"] is not None and banded["contact"] is not None
    assert abs(banded["tau_max"] - alone["tau_max"]) < 1e-6 * alone["tau_max"]
    # an empty band is no band, bit for bit
    m_f, m_p = union(sphere(2, [0, 0, 0], 1.0), sphere(2, [4.0, 0, 0], 1.0))
    half = len(m_p) // 2
    v = np.zeros_like(m_p)
    v[:half] = [1.0, 0.0, 0.0]
    v[half:] = [-1.0, 0.0, 0.0]
    x = fc(m_f, m_p, v, 3.0, l)
    y = fc(m_f, m_p, v, 3.0, l, band=np.zeros(len(m_p), dtype=np.uint8))
    assert x == y and x["contact"] is not None and x["exempt_pairs"] == 0
