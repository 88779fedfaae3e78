"""Pins the FP64 oracle's BVH, face tree, Barnes–Hut lists and Barnes–Hut energy / gradient (CPU).

Following crates/ddg-repulsive-surfaces/tests/barnes_hut.rs and RC/src/spatial/bvh.rs:
* BVH invariants: leaves ≤ leaf size, children partition the parent's range, each split keeps
  ≥ ⌈n/4⌉ on both sides, and every split position minimizes diam²(prefix) + diam²(suffix) over
  the allowed window along the depth's axis, first minimum on ties (checked by brute force);
* cluster data equal direct sums (barnes_hut.rs:253);
* the lists partition every ordered pair (S, T), S ≠ T, exactly once (barnes_hut.rs:214);
* θ = 0 equals the exact energy and gradient (barnes_hut.rs:112); central differences with
  frozen lists (barnes_hut.rs:147); refit equals a fresh evaluation at the same tree;
* the authors' program's Barnes–Hut value within 1e-2 (reference_fixtures.rs:129).
"""
import numpy as np
import pytest

import oracle
from golden import fixtures as fx


@pytest.fixture(scope="module")
def mesh():
    from conftest import load_package

    g = load_package().generate
    p, f = g.torus(16, 10, 1.0, 0.4)
    return g.iso_jitter(p, 0.01, 7), f


def face_data(p, f):
    a0, a1, a2 = p[f[:, 0]], p[f[:, 1]], p[f[:, 2]]
    n = np.cross(a1 - a0, a2 - a0)
    ln = np.sqrt((n * n).sum(axis=1))
    x = ((a0 + a1) + a2) * (1.0 / 3.0)
    lo = np.minimum(np.minimum(a0, a1), a2)
    hi = np.maximum(np.maximum(a0, a1), a2)
    return 0.5 * ln, n / ln[:, None], x, lo, hi


def diam2(lo, hi):
    w = hi - lo
    return (w[0] * w[0] + w[1] * w[1]) + w[2] * w[2]


def test_bvh_structure_and_split_rule(mesh):
    p, f = mesh
    t = oracle.BhTree(f, p, theta=0.5, leaf_size=8).tree()
    area, nrm, x, lo, hi = face_data(p, f)
    order, start, end, right = t["order"], t["start"], t["end"], t["right"]
    nn = len(start)
    assert sorted(order.tolist()) == list(range(len(f)))
    depth = np.zeros(nn, dtype=int)
    for i in range(nn):
        cnt = end[i] - start[i]
        if right[i] == 0:
            assert 1 <= cnt <= 8
            continue
        l, r = i + 1, right[i]
        depth[l] = depth[r] = depth[i] + 1
        assert start[l] == start[i] and end[r] == end[i] and end[l] == start[r]
        k = end[l] - start[l]
        low = min(max(int(np.ceil(0.25 * cnt)), 1), cnt // 2)
        assert low <= k <= cnt - low
        axis = depth[i] % 3
        seg = np.array(sorted(order[start[i]:end[i]].tolist(), key=lambda e: (x[e, axis], e)))
        # the left child holds the first k elements of the node sorted along the depth's axis
        assert set(order[start[l]:end[l]].tolist()) == set(seg[:k].tolist())
        best, best_k = np.inf, cnt // 2
        for kk in range(low, cnt - low + 1):
            pl, ph = lo[seg[:kk]].min(axis=0), hi[seg[:kk]].max(axis=0)
            sl, sh = lo[seg[kk:]].min(axis=0), hi[seg[kk:]].max(axis=0)
            c = diam2(pl, ph) + diam2(sl, sh)
            if c < best:
                best, best_k = c, kk
        assert k == best_k


def test_cluster_data_are_direct_sums(mesh):
    p, f = mesh
    t = oracle.BhTree(f, p, theta=0.5).tree()
    area, nrm, x, lo, hi = face_data(p, f)
    cl = t["clusters"]
    for i in range(len(t["start"])):
        seg = t["order"][t["start"][i]:t["end"][i]]
        a = area[seg].sum()
        assert abs(cl[i, 0] - a) <= 1e-14 * a
        c = (area[seg, None] * x[seg]).sum(axis=0) / a
        assert np.abs(cl[i, 1:4] - c).max() <= 1e-13
        P = np.einsum("s,si,sj->ij", area[seg], nrm[seg], nrm[seg]) / a
        assert np.abs(cl[i, 4:10] - P[np.triu_indices(3)]).max() <= 1e-13
        assert np.array_equal(cl[i, 10:13], lo[seg].min(axis=0)) and np.array_equal(cl[i, 13:16], hi[seg].max(axis=0))
        far = np.maximum(cl[i, 13:16] - c, c - cl[i, 10:13])
        assert abs(cl[i, 16] - np.sqrt((far * far).sum())) <= 1e-13


@pytest.mark.parametrize("theta", [0.25, 0.5, 1.0])
def test_lists_partition_every_pair_once(mesh, theta):
    p, f = mesh
    bh = oracle.BhTree(f, p, theta=theta)
    t, L = bh.tree(), bh.lists()
    nf = len(f)
    leaves = [i for i in range(len(t["start"])) if t["right"][i] == 0]
    cover = np.zeros((nf, nf), dtype=int)
    for k, l in enumerate(leaves):
        S = t["order"][t["start"][l]:t["end"][l]]
        for c in L["lumped"][L["lumped_off"][k]:L["lumped_off"][k + 1]]:
            T = t["order"][t["start"][c]:t["end"][c]]
            cover[np.ix_(S, T)] += 1
        for c in L["exact"][L["exact_off"][k]:L["exact_off"][k + 1]]:
            T = t["order"][t["start"][c]:t["end"][c]]
            cover[np.ix_(S, T)] += 1
    off = ~np.eye(nf, dtype=bool)
    assert np.all(cover[off] == 1)
    assert np.all(cover.diagonal() == 1), "a face meets itself once (its own leaf, skipped in the sum)"


def test_theta_zero_is_exact(mesh):
    p, f = mesh
    bh = oracle.BhTree(f, p, theta=0.0)
    e, d, _ = bh.energy_and_differential()
    e0, d0, _ = oracle.exact(f, p)
    assert abs(e - e0) <= 1e-13 * e0
    assert np.abs(d - d0).max() <= 1e-12 * np.abs(d0).max()
    assert bh.energy() == e


def test_frozen_list_gradient_by_central_differences(mesh):
    p, f = mesh
    bh = oracle.BhTree(f, p, theta=0.5)
    _, g, _ = bh.energy_and_differential()
    edges = np.concatenate([f[:, [0, 1]], f[:, [1, 2]], f[:, [2, 0]]])
    h = 1e-6 * np.linalg.norm(p[edges[:, 0]] - p[edges[:, 1]], axis=1).mean()
    rng = np.random.default_rng(1)
    idx = rng.choice(len(p), 12, replace=False)
    err, nrm = 0.0, 0.0
    for v in idx:
        for k in range(3):
            q = p.copy()
            q[v, k] += h
            bh.refit(q)
            ep = bh.energy()
            q[v, k] -= 2 * h
            bh.refit(q)
            em = bh.energy()
            fd = (ep - em) / (2 * h)
            err += (fd - g[v, k]) ** 2
            nrm += g[v, k] ** 2
    assert np.sqrt(err / nrm) < 1e-6
    bh.refit(p)
    e1, g1, _ = bh.energy_and_differential()
    bh2 = oracle.BhTree(f, p, theta=0.5)
    e2, g2, _ = bh2.energy_and_differential()
    assert e1 == e2 and np.array_equal(g1, g2), "refit back to the build positions reproduces the build"


def test_error_decreases_like_theta_squared(mesh):
    p, f = mesh
    e0, _, _ = oracle.exact(f, p)
    errs = [abs(oracle.BhTree(f, p, theta=t).energy() - e0) / e0 for t in (1.0, 0.5, 0.25)]
    assert errs[1] < 5e-2 and errs[2] < errs[0]


@pytest.mark.parametrize("name", fx.AUTHORS_NAMES)
def test_authors_program_barnes_hut(name):
    fxd = fx.authors(name)
    kv = fxd["kv"]
    bh = oracle.BhTree(fxd["faces"], fxd["positions"], theta=kv["theta"], leaf_size=8)
    e, d, _ = bh.energy_and_differential(kv["alpha"], kv["beta"])
    assert abs(e - kv["tpe_barnes_hut"]) / kv["tpe_barnes_hut"] < 1e-2
    assert np.linalg.norm(d - fxd["d_bh"]) / np.linalg.norm(fxd["d_bh"]) < 5e-2
