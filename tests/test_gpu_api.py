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
    api = pkg.api.RepulsiveSurfaceSession(p1.reshape(-1), f1.reshape(-1), device="rocm", theta=0.5)
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
    api = pkg.api.RepulsiveSurfaceSession(p, f, device="rocm", constraints=cons, energy_eval="exact", chi=0.0)
    native = pkg.Session(pkg.Surface(ctx, f, p), pkg.StepOptions(energy_exact=True, saddle=pkg.SaddleOptions(chi=0.0),
                                                                 constraints=(("area", None), ("volume", None)),
                                                                 ccd=pkg.CcdOptions()))
    for _ in range(2):
        a, n = api.step(), native.step()
        assert a["status"] == "accepted" and a["energy_after"] == n.energy_after
        assert a["newton_steps"] == n.newton_rounds >= 1 and len(a["phi"]) == 5
        print("phi after", a["phi"], "rounds", a["newton_steps"])
    assert np.array_equal(api.positions.reshape(-1, 3), native.positions())


def test_surface_energy_and_gpu_info(pkg, ctx):
    p, f = pkg.generate.config_icosphere_l3()
    e, d, rep = pkg.api.surface_energy(p, f, energy_eval="exact")
    e0, d0 = pkg.exact_energy_and_differential(pkg.Surface(ctx, f, p))
    assert e == e0 and np.array_equal(d, d0.reshape(-1)) and rep["pairs_exact"] == len(f) * (len(f) - 1)
    info = pkg.api.gpu_info()
    print(info)
    assert info["available"] and info["devices"][0]["ordinal"] == 0
