"""The reference-shaped Python API on the device (GPU): ``RepulsiveSurfaceSession`` with
``device="rocm"`` takes the same steps as ``Session`` with the same options, ``run`` stops at
``max_steps``, ``surface_energy`` equals the exact energy entry point, ``gpu_info`` lists the
device, and an OBJ written and read back drives the same session."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_session_api_matches_native(pkg, ctx, tmp_path):
    p, f = pkg.generate.config_icosphere_l3()
    path = tmp_path / "l3.obj"
    pkg.io.write_obj(path, p, f)
    p1, f1 = pkg.io.read_obj(path)
    api = pkg.api.RepulsiveSurfaceSession(p1.reshape(-1), f1.reshape(-1), device="rocm", theta=0.5, remesh=False)
    native = pkg.Session(pkg.Surface(ctx, f, p), pkg.StepOptions(ccd=pkg.CcdOptions()))  # the API's default CCD
    for _ in range(2):
        a, n = api.step(), native.step()
        assert a["status"] == "accepted" and a["energy_before"] == n.energy_before and a["tau"] == n.tau
        assert a["tau_ccd"] == n.tau_ccd and n.ccd is not None
        assert a["linear"]["converged"] and a["linear"]["iterations"] == n.gmres.iterations
    assert np.array_equal(api.positions.reshape(-1, 3), native.positions())
    out = api.run(max_steps=2)
    assert out["end"] == "max_steps" and out["steps_taken"] == 2 and api.iteration == 4
    snap = api.snapshot()
    assert snap["iteration"] == 4 and snap["report"]["iteration"] == 4
    api.reset()
    assert api.iteration == 0 and np.array_equal(api.positions.reshape(-1, 3), p1)


def test_session_api_constraints(pkg, ctx):
    """The reference's ``constraints`` list (area and volume held) through the API: the same steps
    as the native session with the same rows, Φ reported before and after the projection."""
    p, f = pkg.generate.config_icosphere_l3()
    cons = [{"type": "barycenter"}, {"type": "area"}, {"type": "volume"}]
    api = pkg.api.RepulsiveSurfaceSession(p, f, device="rocm", constraints=cons, energy_eval="exact", chi=0.0,
                                          remesh=False)
    native = pkg.Session(pkg.Surface(ctx, f, p), pkg.StepOptions(energy_exact=True, saddle=pkg.SaddleOptions(chi=0.0),
                                                                 constraints=(("area", None), ("volume", None)),
                                                                 ccd=pkg.CcdOptions()))
    for _ in range(2):
        a, n = api.step(), native.step()
        assert a["status"] == "accepted" and a["energy_after"] == n.energy_after
        assert a["newton_steps"] == n.newton_rounds >= 1 and len(a["phi"]) == 5
        print("phi after", a["phi"], "rounds", a["newton_steps"])
    assert np.array_equal(api.positions.reshape(-1, 3), native.positions())


def test_session_api_remeshes_after_each_step(pkg, ctx):
    """Remeshing on (the reference's default): the first step's remeshing equals the oracle's pass
    of the same projected positions (the native session's), with the collapse guard at the line
    search's CCD gap and the band of the faces that intersect at the step's start; the run then
    continues on the remeshed mesh with the initial ℓ̄₀ and the face origins composed."""
    import oracle

    base, f = pkg.generate.icosphere(3)
    f = np.asarray(f, np.uint32)
    p = np.asarray(base) * np.array([1.6, 1.0, 0.7])
    api = pkg.api.RepulsiveSurfaceSession(p, f, device="rocm")
    native = pkg.Session(pkg.Surface(ctx, f, p), pkg.StepOptions(ccd=pkg.CcdOptions()))
    l0 = native.state()["mean_edge_length0"]
    band = pkg.intersection_band(native.surface, p)[0]
    a, n = api.step(), native.step()
    assert a["energy_before"] == n.energy_before and a["tau"] == n.tau
    ref, rref = oracle.remesh(f, native.positions(), l0, guard_gap=0.05 * l0, band=band)
    rep = a["remesh"]
    for k in ("splits", "collapses", "flips", "smoothing_moves", "refused_collapses", "refused_flips"):
        assert rep[k] == rref[k], k
    assert rep["splits"] + rep["collapses"] + rep["flips"] > 0
    np.testing.assert_array_equal(api.faces.reshape(-1, 3), ref["faces"])
    assert float(np.abs(api.positions.reshape(-1, 3) - ref["positions"]).max()) <= 1e-12
    assert a["n_vertices"] == len(ref["positions"]) and a["n_faces"] == len(ref["faces"])
    np.testing.assert_array_equal(api.snapshot()["face_origin"], ref["face_parents"])
    b = api.step()
    assert b["status"] == "accepted" and np.isfinite(b["energy_after"])
    assert api.snapshot()["mean_edge_length0"] == l0
    native.close()


def test_surface_energy_and_gpu_info(pkg, ctx):
    p, f = pkg.generate.config_icosphere_l3()
    e, d, rep = pkg.api.surface_energy(p, f, energy_eval="exact")
    e0, d0 = pkg.exact_energy_and_differential(pkg.Surface(ctx, f, p))
    assert e == e0 and np.array_equal(d, d0.reshape(-1)) and rep["pairs_exact"] == len(f) * (len(f) - 1)
    info = pkg.api.gpu_info()
    print(info)
    assert info["available"] and info["devices"][0]["ordinal"] == 0
