"""Hierarchical products on the device against the FP64 oracle (GPU).

B̃, B̃⁰ (face pairs, χ = 0.5, leaf 8, leaf_pairs 64) and L̃^V_{2−s} (vertex pairs) in FP32
(near and far values evaluated in FP64, stored FP32; FP32 accumulation) against the oracle's FP64
product on the same block tree: relative error ≤ 1e-5 in the 2-norm; block counts equal.

The block cluster trees built on the device equal the oracle's CPU build bit for bit: the BVH
(order, start, end, right) and the admissible and inadmissible block lists per row node in
creation order (RS/src/spatial/block_tree.rs:129-209), for the face tree and the vertex tree, on
the level-3 icosphere, the 16,384-face torus and the level-6 bumpy icosphere.
"""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

TOL = 1e-5


BLOCK_KEYS = ("order", "start", "end", "right", "far_ptr", "far_col", "near_ptr", "near_col")


def assert_same_blocks(dev, orc, what):
    for k in BLOCK_KEYS:
        assert dev[k].shape == orc[k].shape and np.array_equal(dev[k], orc[k]), (what, k, dev[k].shape, orc[k].shape)
    print(f"{what}: {len(dev['order'])} elements, {len(dev['start'])} nodes, {len(dev['far_col'])} admissible and "
          f"{len(dev['near_col'])} inadmissible blocks, equal to the CPU build")


def meshes(pkg):
    yield "icosphere_l3", *pkg.generate.config_icosphere_l3()
    yield "torus_16k", *pkg.generate.config_torus_16k()


@pytest.mark.parametrize("which", [0, 1])
def test_face_metric(pkg, ctx, which):
    name, p, f = list(meshes(pkg))[which]
    s = pkg.Surface(ctx, f, p)
    fm = pkg.FaceMetricTree(s)
    om = oracle.FaceMetric(f, p, chi=0.5)
    st, so = fm.stats(), om.stats()
    for k in ("admissible_blocks", "inadmissible_blocks", "admissible_pairs", "near_entries", "nodes", "depth", "near_total"):
        assert st[k] == so[k], (k, st[k], so[k])
    assert_same_blocks(fm.blocks(), om.blocks(), f"{name} face tree")
    rng = np.random.default_rng(11)
    x = rng.standard_normal((3, 3 * len(p)))
    for w_hi, w_lo in [(1.0, 1.0), (1.0, 0.0), (0.0, 1.0)]:
        y = fm.apply_batch(x, w_hi, w_lo)
        y0 = om.apply(x, w_hi, w_lo)
        err = np.linalg.norm(y - y0) / np.linalg.norm(y0)
        print(f"{name} ({w_hi},{w_lo}): rel {err:.2e}")
        assert err <= TOL
    ones = np.zeros(3 * len(p))
    ones[1::3] = 1.0
    y = fm.apply_batch(ones)
    assert np.abs(y).max() <= 1e-5 * np.abs(fm.apply_batch(x[0])).max()


@pytest.mark.parametrize("which", [0, 1])
def test_vertex_fractional(pkg, ctx, which):
    name, p, f = list(meshes(pkg))[which]
    s = pkg.Surface(ctx, f, p)
    tau = 2.0 - (12.0 - 2.0) / 6.0
    vf = pkg.VertexFractionalTree(s, tau)
    mass = oracle.vertex_areas(f, p)
    ov = oracle.VertexFractional(p, mass, tau, chi=0.5)
    assert_same_blocks(vf.blocks(), ov.blocks(), f"{name} vertex tree")
    rng = np.random.default_rng(12)
    x = rng.standard_normal((2, 3 * len(p)))
    y, y0 = vf.apply_batch(x), ov.apply(x)
    err = np.linalg.norm(y - y0) / np.linalg.norm(y0)
    print(f"{name} L^V: rel {err:.2e}")
    assert err <= TOL


def test_face_metric_l6(pkg, ctx):
    p, f = pkg.generate.config_bumpy_l6()
    s = pkg.Surface(ctx, f, p)
    fm = pkg.FaceMetricTree(s)
    om = oracle.FaceMetric(f, p, chi=0.5)
    assert fm.stats() == om.stats()
    assert_same_blocks(fm.blocks(), om.blocks(), "bumpy_l6 face tree")
    rng = np.random.default_rng(13)
    x = rng.standard_normal(3 * len(p))
    y, y0 = fm.apply_batch(x), om.apply(x)
    err = np.linalg.norm(y - y0) / np.linalg.norm(y0)
    print(f"L6 metric: rel {err:.2e}; stats {fm.stats()}")
    assert err <= TOL


@pytest.mark.parametrize("which", [0, 1])
def test_block_build_device_equals_host(pkg, ctx, which, monkeypatch):
    """The device build of the block structures and panels equals the host planner's bitwise."""
    name, p, f = list(meshes(pkg))[which]
    s = pkg.Surface(ctx, f, p)
    tau = 2.0 - (12.0 - 2.0) / 6.0
    rng = np.random.default_rng(14)
    x = rng.standard_normal((2, 3 * len(p)))
    fm_d, vf_d = pkg.FaceMetricTree(s), pkg.VertexFractionalTree(s, tau)
    monkeypatch.setenv("DDGRS_BLOCKS_ON_HOST", "1")
    fm_h, vf_h = pkg.FaceMetricTree(s), pkg.VertexFractionalTree(s, tau)
    monkeypatch.delenv("DDGRS_BLOCKS_ON_HOST")
    assert fm_d.stats() == fm_h.stats()
    assert np.array_equal(fm_d.apply_batch(x, 1.0, 0.3), fm_h.apply_batch(x, 1.0, 0.3))
    assert np.array_equal(vf_d.apply_batch(x), vf_h.apply_batch(x))


@pytest.mark.parametrize("seg", ["64", "192"])
def test_panel_segments(pkg, ctx, seg, monkeypatch):
    """Near-field panels split into column segments (panels of several segments finished from
    their segments' stored sums) give the product of unsplit panels to FP32 summation order, on
    both operators, and still match the oracle."""
    p, f = pkg.generate.config_torus_16k()
    s = pkg.Surface(ctx, f, p)
    tau = 2.0 - (12.0 - 2.0) / 6.0
    rng = np.random.default_rng(15)
    x = rng.standard_normal((2, 3 * len(p)))
    fm_a, vf_a = pkg.FaceMetricTree(s), pkg.VertexFractionalTree(s, tau)
    monkeypatch.setenv("DDGRS_SEGMENT_COLUMNS", seg)
    fm_b, vf_b = pkg.FaceMetricTree(s), pkg.VertexFractionalTree(s, tau)
    monkeypatch.setenv("DDGRS_BLOCKS_ON_HOST", "1")
    fm_h = pkg.FaceMetricTree(s)
    monkeypatch.delenv("DDGRS_BLOCKS_ON_HOST")
    monkeypatch.delenv("DDGRS_SEGMENT_COLUMNS")
    ya, yb = fm_a.apply_batch(x, 1.0, 0.3), fm_b.apply_batch(x, 1.0, 0.3)
    assert np.array_equal(yb, fm_h.apply_batch(x, 1.0, 0.3))
    assert np.linalg.norm(ya - yb) / np.linalg.norm(ya) <= 1e-6
    y0 = oracle.FaceMetric(f, p, chi=0.5).apply(x, 1.0, 0.3)
    assert np.linalg.norm(yb - y0) / np.linalg.norm(y0) <= TOL
    va, vb = vf_a.apply_batch(x), vf_b.apply_batch(x)
    assert np.linalg.norm(va - vb) / np.linalg.norm(va) <= 1e-6
    # applied twice: the per-panel arrival counts are back at zero after each product
    assert np.array_equal(fm_b.apply_batch(x, 1.0, 0.3), yb)
    assert np.array_equal(vf_b.apply_batch(x), vb)


def test_panel_segments_invalid(pkg, ctx, monkeypatch):
    p, f = pkg.generate.config_icosphere_l3()
    s = pkg.Surface(ctx, f, p)
    monkeypatch.setenv("DDGRS_SEGMENT_COLUMNS", "100")
    with pytest.raises(Exception, match="multiple of 64"):
        pkg.FaceMetricTree(s)


@pytest.mark.parametrize("which", [0, 1])
def test_extended_products(pkg, ctx, which, monkeypatch):
    """The extended near field (FP64 differences and sums of the same FP32 values) gives the FP32
    product to its rounding and is at least as close to the FP64 oracle, for B̃/B̃⁰ and L̃^V; on a
    smooth input, where the FP32 sums cancel, it is much closer."""
    name, p, f = list(meshes(pkg))[which]
    s = pkg.Surface(ctx, f, p)
    tau = 2.0 - (12.0 - 2.0) / 6.0
    fm, vf = pkg.FaceMetricTree(s), pkg.VertexFractionalTree(s, tau)
    om = oracle.FaceMetric(f, p, chi=0.5)
    ov = oracle.VertexFractional(p, oracle.vertex_areas(f, p), tau, chi=0.5)
    rng = np.random.default_rng(16)
    smooth = np.concatenate([p, p * p], axis=1)[:, :3].reshape(-1)  # a smooth field: positions and squares
    x = np.stack([rng.standard_normal(3 * len(p)), smooth])
    y32, v32 = fm.apply_batch(x, 1.0, 0.3), vf.apply_batch(x)
    monkeypatch.setenv("DDGRS_PRODUCTS_EXTENDED", "1")
    y64, v64 = fm.apply_batch(x, 1.0, 0.3), vf.apply_batch(x)
    monkeypatch.delenv("DDGRS_PRODUCTS_EXTENDED")
    y0, v0 = om.apply(x, 1.0, 0.3), ov.apply(x)
    for k in range(2):
        e32 = np.linalg.norm(y32[k] - y0[k]) / np.linalg.norm(y0[k])
        e64 = np.linalg.norm(y64[k] - y0[k]) / np.linalg.norm(y0[k])
        f32 = np.linalg.norm(v32[k] - v0[k]) / np.linalg.norm(v0[k])
        f64 = np.linalg.norm(v64[k] - v0[k]) / np.linalg.norm(v0[k])
        print(f"{name} input {k}: metric FP32 {e32:.2e} extended {e64:.2e}; L^V FP32 {f32:.2e} extended {f64:.2e}")
        assert e64 <= TOL and f64 <= TOL
        assert e64 <= 1.01 * e32 + 1e-12 and f64 <= 1.01 * f32 + 1e-12
        assert np.linalg.norm(y64[k] - y32[k]) / np.linalg.norm(y64[k]) <= 1e-5


def test_vertex_tree_blocks_l6(pkg, ctx):
    p, f = pkg.generate.config_bumpy_l6()
    s = pkg.Surface(ctx, f, p)
    tau = 2.0 - (12.0 - 2.0) / 6.0
    vf = pkg.VertexFractionalTree(s, tau)
    ov = oracle.VertexFractional(p, oracle.vertex_areas(f, p), tau, chi=0.5)
    assert_same_blocks(vf.blocks(), ov.blocks(), "bumpy_l6 vertex tree")


def test_host_planned_blocks_export_the_same_lists(pkg, ctx, monkeypatch):
    """The lists read back from a host-planned build equal those of the device build."""
    p, f = pkg.generate.config_torus_16k()
    s = pkg.Surface(ctx, f, p)
    dev = pkg.FaceMetricTree(s).blocks()
    monkeypatch.setenv("DDGRS_BLOCKS_ON_HOST", "1")
    host = pkg.FaceMetricTree(s).blocks()
    monkeypatch.delenv("DDGRS_BLOCKS_ON_HOST")
    assert_same_blocks(dev, host, "torus_16k host plan")
