"""Pins the FP64 oracle's hierarchical metric products (CPU only).

* χ = 0 (every block exact) reproduces the Lean golden dense metric ``rs-*-metric`` (B, B⁰, A)
  to 1e-12 (crates/ddg-repulsive-surfaces/tests/hierarchical.rs:169, oracle.rs:41);
* χ = 0.5 stays within 5 % of χ = 0 (hierarchical.rs:192);
* the null space (B̃⁰ 1 = 0 exactly through H̃a, hierarchical.rs:4-16) and symmetry;
* D_f and the face average of the Lean golden ``rs-{tetra,icosa,patch,mobius}-df`` fixtures to
  1e-12 normwise (tests/oracle.rs:262 ``run_df``, metric/mod.rs:120-225), through the same
  ``df_coefficients`` the metric products use; plus D_f 1 = 0 and D_f of a linear function
  equal to its tangential gradient;
* the vertex-pair L^V_τ at χ = 0 equals the brute-force formula
  Σ_{v≠w} (u_v − u_w)² a_v a_w |f_v − f_w|^{−(2τ+2)} (dense.rs:479).
"""
import numpy as np
import pytest

import oracle
from golden import fixtures as fx


def dense_from_apply(fm, nv, w_hi, w_lo):
    X = np.zeros((nv, 3 * nv))
    for j in range(nv):
        X[j, 3 * j] = 1.0
    return fm.apply(X, w_hi, w_lo)[:, 0::3].T


@pytest.mark.parametrize("name", fx.cases("metric"))
def test_lean_metric_fixtures_at_chi_zero(name):
    c, o = fx.read_case(name), fx.read_out(name)
    a, b = c["exponents"]
    w_tp, c_lo, w_w = c.get("weights", (1.0, 1.0, 0.0))
    nv = len(c["positions"])
    fm = oracle.FaceMetric(c["faces"], c["positions"], a, b, chi=0.0)
    B = dense_from_apply(fm, nv, 1.0, 0.0)
    B0 = dense_from_apply(fm, nv, 0.0, 1.0)
    Bref, B0ref = fx.matrix(o, "B"), fx.matrix(o, "B0")
    assert np.abs(B - Bref).max() <= 1e-12 * np.abs(Bref).max()
    assert np.abs(B0 - B0ref).max() <= 1e-12 * np.abs(B0ref).max()
    if w_w == 0.0:
        A = dense_from_apply(fm, nv, w_tp, w_tp * c_lo)
        Aref = fx.matrix(o, "A")
        assert np.abs(A - Aref).max() <= 1e-12 * np.abs(Aref).max()


@pytest.fixture(scope="module")
def torus():
    from conftest import load_package

    g = load_package().generate
    p, f = g.torus(24, 12, 1.0, 0.4)
    return g.iso_jitter(p, 0.01, 5), f


def test_chi_half_within_five_percent(torus):
    p, f = torus
    exact = oracle.FaceMetric(f, p, chi=0.0)
    approx = oracle.FaceMetric(f, p, chi=0.5)
    assert approx.stats()["admissible_blocks"] > 0
    rng = np.random.default_rng(3)
    x = rng.standard_normal((2, 3 * len(p)))
    y0, y1 = exact.apply(x), approx.apply(x)
    err = np.linalg.norm(y1 - y0) / np.linalg.norm(y0)
    assert err < 5e-2, err


def test_null_space_and_symmetry(torus):
    p, f = torus
    fm = oracle.FaceMetric(f, p, chi=0.5)
    nv = len(p)
    ones = np.zeros(3 * nv)
    ones[0::3] = 1.0
    y_lo = fm.apply(ones, 0.0, 1.0)
    assert np.all(y_lo == 0.0)  # H̃a computed with the same H̃: bitwise zero
    y_hi = fm.apply(ones, 1.0, 0.0)
    rng = np.random.default_rng(4)
    u, v = rng.standard_normal((2, 3 * nv))
    scale = np.abs(fm.apply(u)).max()
    assert np.abs(y_hi).max() <= 1e-12 * scale
    assert abs(u @ fm.apply(v) - v @ fm.apply(u)) <= 1e-12 * np.abs(u) @ np.abs(fm.apply(v))


def test_vertex_fractional_brute_force(torus):
    p, f = torus
    nv = len(p)
    mass = oracle.vertex_areas(f, p)
    s = (12.0 - 2.0) / 6.0
    tau = 2.0 - s
    vf = oracle.VertexFractional(p, mass, tau, chi=0.0)
    rng = np.random.default_rng(5)
    x = rng.standard_normal(3 * nv)
    y = vf.apply(x)
    d = p[:, None, :] - p[None, :, :]
    r2 = (d * d).sum(axis=2)
    np.fill_diagonal(r2, 1.0)
    H = r2 ** (-(tau + 1.0))
    np.fill_diagonal(H, 0.0)
    # L = 2 (diag(a ⊙ H a) − diag(a) H diag(a)) per axis
    L = 2.0 * (np.diag(mass * (H @ mass)) - mass[:, None] * H * mass[None, :])
    want = np.stack([L @ x[k::3] for k in range(3)], axis=1).reshape(-1)
    assert np.abs(y - want).max() <= 1e-11 * np.abs(want).max()
    approx = oracle.VertexFractional(p, mass, tau, chi=0.5)
    assert np.linalg.norm(approx.apply(x) - want) / np.linalg.norm(want) < 5e-2


@pytest.mark.parametrize("name", ["rs-tetra-df", "rs-icosa-df", "rs-patch-df", "rs-mobius-df"])
def test_lean_df_fixtures(name):
    c, o = fx.read_case(name), fx.read_out(name)
    got, want = [], []
    for k, f, vals in o["df"]:
        avg, grad = oracle.face_operators(c["faces"], c["positions"], c["u"][k])
        got += [avg[f], *grad[f]]
        want += vals
    got, want = np.array(got), np.array(want)
    assert len(want) >= 8
    assert np.abs(got - want).max() <= 1e-12 * np.abs(want).max()


def test_df_identities():
    """D_f 1 = 0, and for u = ⟨a, f⟩ (linear) D_f u = a − ⟨a, N⟩ N on every face."""
    c = fx.read_case("rs-icosa-df")
    p, f = c["positions"], c["faces"].astype(np.int64)
    _, g1 = oracle.face_operators(f, p, np.ones(len(p)))
    assert np.abs(g1).max() <= 1e-14
    a = np.array([0.3, -1.1, 0.7])
    _, g = oracle.face_operators(f, p, p @ a)
    n = np.cross(p[f[:, 1]] - p[f[:, 0]], p[f[:, 2]] - p[f[:, 0]])
    n /= np.linalg.norm(n, axis=1)[:, None]
    tang = a[None, :] - (n @ a)[:, None] * n
    assert np.abs(g - tang).max() <= 1e-13


def test_block_lists_partition_every_pair():
    """The oracle's admissible and inadmissible blocks cover every ordered element pair exactly
    once (block_tree.rs: the level-by-level split of (root, root)), admissible blocks are never
    diagonal, and the counts match the stats."""
    c = fx.read_case("rs-torus-metric")
    p, f = c["positions"], c["faces"]
    fm = oracle.FaceMetric(f, p, chi=0.5, leaf_size=2, leaf_pairs=4)
    b = fm.blocks()
    n = len(b["order"])
    cover = np.zeros((n, n), dtype=np.int64)
    for key in ("far", "near"):
        ptr, col = b[key + "_ptr"], b[key + "_col"]
        for a in range(len(b["start"])):
            for q in range(ptr[a], ptr[a + 1]):
                c_ = col[q]
                if key == "far":
                    assert a != c_
                cover[b["start"][a]:b["end"][a], b["start"][c_]:b["end"][c_]] += 1
    assert np.all(cover == 1)
    st = fm.stats()
    assert len(b["far_col"]) == st["admissible_blocks"] and len(b["near_col"]) == st["inadmissible_blocks"]
