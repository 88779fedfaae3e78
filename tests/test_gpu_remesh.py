"""The device remeshing pass (ddgrs_remesh, ddgrs_remesh_op) against the FP64 oracle and the
reference's Lean single-operation fixtures.

* The 16 Lean ``remesh-ops`` fixtures replayed on the device exactly as on the oracle
  (tests/test_oracle_remesh.py): refusal kinds, faces, positions bit for bit, parents, vertex map.
* Whole passes on closed, bounded, non-orientable and two-component meshes, a stretched
  icosphere, and the bumpy icosphere at L5 (20,480 faces) after descent-like motion: the device
  result equals the oracle's — the same faces, face parents, vertex map, pins and report counts,
  positions bit for bit where no smoothing ran and within 1e-12 of the bounding box after
  smoothing (atan2 in the vertex normal is the device library's, not glibc's; every other
  operation is the same FP64 expression).  Each pass starts both sides from the same input.
* The collapse guard (two sheets within the gap): the same refusals as the oracle's grid guard.
"""
import numpy as np
import pytest

import oracle
from conftest import load_package
from golden import fixtures as fx
from test_oracle_remesh import OPS, _mean_edge, grid, moebius, named_meshes

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def gctx(pkg):
    c = pkg.Context(0)
    yield c
    c.close()


@pytest.mark.parametrize("name", OPS)
def test_remesh_ops_fixture_on_device(pkg, gctx, name):
    c = fx.read_case(name)
    o = fx.read_out(name)
    op, a, b = c["op"]
    pins = c.get("pinned", [])
    leg = c.get("legality", set())
    print(f"remesh op fixture {name}: {op} {a} {b}", flush=True)
    refusal, reordered, r = pkg.remesh_op(gctx, c["faces"], c["positions"], op, a, b, pinned=pins,
                                          delaunay="delaunay" in leg, foldover="foldover" in leg)
    if o["status"] == "error":
        kind = o["error"][0]
        if kind == "InconsistentOrientation":
            assert refusal is None and reordered
            assert len(r.faces) == len(c["faces"])
            return
        expect = {"NonManifoldResult": ("RepeatedFace", "NonManifoldResult"),
                  "FlipNonManifold": ("ExistingEdge",)}.get(kind, (kind,))
        assert refusal in expect, (name, refusal, kind)
        return
    assert refusal is None and not reordered
    np.testing.assert_array_equal(r.faces, np.array(o["faces"], dtype=np.uint32))
    assert r.positions.tobytes() == np.array(o["positions"], dtype=np.float64).tobytes()
    np.testing.assert_array_equal(r.face_parents, o["parent"])
    vm = [list(r.vertex_map.entry(v)[0]) for v in range(len(r.positions))]
    assert vm == o["vmap"]


def _compare(dev, rep, ref, rref, tol, name):
    np.testing.assert_array_equal(dev.faces, ref["faces"], err_msg=name)
    np.testing.assert_array_equal(dev.face_parents, ref["face_parents"], err_msg=name)
    np.testing.assert_array_equal(dev.vertex_map.offsets, ref["map_offsets"], err_msg=name)
    np.testing.assert_array_equal(dev.vertex_map.sources, ref["map_sources"], err_msg=name)
    assert dev.vertex_map.weights.tobytes() == ref["map_weights"].tobytes(), name
    np.testing.assert_array_equal(dev.pinned, ref["pinned"], err_msg=name)
    scale = float(np.ptp(ref["positions"], axis=0).max())
    err = float(np.abs(dev.positions - ref["positions"]).max())
    if tol == 0.0:
        assert dev.positions.tobytes() == ref["positions"].tobytes(), (name, err)
    else:
        assert err <= tol * scale, (name, err / scale)
    for k in ("splits", "collapses", "flips", "splits_deferred", "smoothing_moves", "smoothing_halved",
              "smoothing_refused", "smoothing_skipped", "flip_cap_hits", "n_vertices", "n_faces"):
        assert getattr(rep, k) == rref[k], (name, k, getattr(rep, k), rref[k])
    for k in ("refused_splits", "refused_collapses", "refused_flips"):
        assert getattr(rep, k) == rref[k], (name, k)
    return err / scale


def _passes(pkg, gctx, name, faces, pos, l0, npass, pins=(), **kw):
    f, q, pn = np.asarray(faces, np.uint32), np.asarray(pos, np.float64), list(pins)
    worst = 0.0
    no_smooth = "smooth" not in kw.get("enabled", ("smooth",))
    for i in range(npass):
        print(f"remesh pass {name} #{i}: {len(q)} vertices, {len(f)} faces", flush=True)
        ref, rref = oracle.remesh(f, q, l0, pinned=pn, **kw)
        opts = pkg.RemeshOptions(enabled_ops=pkg.EnabledOps(*(k in kw.get("enabled", ("split", "collapse", "flip",
                                                                                         "smooth"))
                                                              for k in ("split", "collapse", "flip", "smooth"))))
        dev, rep = pkg.remesh(gctx, f, q, l0, opts, pinned=pn)
        worst = max(worst, _compare(dev, rep, ref, rref, 0.0 if no_smooth else 1e-12, f"{name} pass {i}"))
        f, q, pn = ref["faces"], ref["positions"], list(ref["pinned"])
    return worst


@pytest.mark.parametrize("idx", range(5))
def test_full_passes_match_oracle(pkg, gctx, idx):
    name, faces, pos = named_meshes()[idx]
    rng = np.random.default_rng(idx)
    p = pos * np.array((3.0, 1.0, 0.5)) + rng.uniform(-0.01, 0.01, pos.shape)
    l0 = _mean_edge(faces, p)
    _passes(pkg, gctx, name, faces, p, l0, 3, pins=[0])


def test_stretched_icosphere_matches_oracle(pkg, gctx):
    gen = pkg.generate
    p, f = gen.icosphere(3)
    f = np.asarray(f, np.uint32)
    base = np.asarray(p)
    l0 = _mean_edge(f, base) * 1.5
    _passes(pkg, gctx, "stretched", f, base * np.array([4.0, 1.0, 0.5]), l0, 4)


def test_without_smoothing_positions_are_bitwise(pkg, gctx):
    gen = pkg.generate
    p, f = gen.icosphere(3)
    q = np.asarray(p) * np.array([2.5, 1.0, 0.4])
    l0 = _mean_edge(np.asarray(f, np.uint32), np.asarray(p))
    _passes(pkg, gctx, "no-smooth", f, q, l0, 2, enabled=("split", "collapse", "flip"))


def test_bumpy_l5_pass_matches_oracle(pkg, gctx):
    gen = pkg.generate
    p, f = gen.icosphere(5)
    p = np.asarray(p)
    f = np.asarray(f, np.uint32)
    l0 = _mean_edge(f, p)
    q = gen.bumps(p, 16, (0.0, 0.3), (0.1, 0.4), 7)
    _passes(pkg, gctx, "bumpy-l5", f, q, l0, 2)


def test_collapse_guard_matches_oracle(pkg, gctx):
    f1, p1 = grid(8)
    top = p1 * np.array([0.45, 1.0, 1.0]) + np.array([0.25, 0.0, 0.02])
    faces = np.vstack([f1, f1 + len(p1)]).astype(np.uint32)
    pos = np.vstack([p1, top])
    l0 = 1.0 / 8
    for gap, band in ((0.03, None), (0.03, np.r_[np.ones(10, np.uint8), np.zeros(len(pos) - 10, np.uint8)]),
                      (0.005, None)):
        ref, rref = oracle.remesh(faces, pos, l0, enabled=("collapse",), guard_gap=gap, band=band)
        dev, rep = pkg.remesh(gctx, faces, pos, l0, pkg.RemeshOptions(enabled_ops=pkg.EnabledOps(False, True, False,
                                                                                               False)),
                              guard_gap=gap, band=band)
        _compare(dev, rep, ref, rref, 0.0, f"guard {gap}")
    assert rref["collapses"] > 0 and rep.refused_collapses["Guard"] == rref["refused_collapses"]["Guard"]


def test_options_and_inputs_are_validated(pkg, gctx):
    f, p = grid(3)
    for kw in ({"split": 0.0}, {"collapse": -1.0}, {"split": 1.0, "collapse": 0.5}, {"rho": 0.0},
               {"flip_cap_per_vertex": 0}, {"max_growth": 0.5}):
        with pytest.raises(pkg.RsError) as e:
            pkg.remesh(gctx, f, p, 0.3, pkg.RemeshOptions(**kw))
        assert e.value.kind == "InvalidOption"
    with pytest.raises(pkg.RsError):
        pkg.remesh(gctx, f, p, 0.0)
    with pytest.raises(pkg.RsError):
        pkg.remesh(gctx, f, p, 0.3, pinned=[len(p)])
    bad = p.copy()
    bad[5] = bad[6]
    with pytest.raises(pkg.RsError) as e:
        pkg.remesh(gctx, f, bad, 0.3)
    assert e.value.kind in ("DegenerateFace", "NonFinitePosition")
