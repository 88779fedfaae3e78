"""OBJ input/output and the Python option layer (CPU: no device call is made).

* ``parse_obj`` follows crates/ddg-core/src/io.rs:166-350: 1-based and negative indices, ``v/vt/vn``
  corners, comments, continued lines, forward references checked at the end; only triangles
  (``PolygonFace``); parse errors name the line.  The authors' OBJ fixtures read back to the
  same arrays as the test reader, and ``obj_to_string`` round-trips exactly.
* The options of ``RepulsiveSurfaceSession`` map to the step options by the reference's names
  (crates/ddg-py/src/repulsive_surfaces.rs:786-1050); choices outside this path are
  ``Unsupported``, unknown names ``InvalidOption``.
"""
import numpy as np
import pytest

from conftest import load_package
from golden import fixtures as fx


def test_parse_obj_rules():
    pkg = load_package()
    text = ("# a comment\nv 0 0 0\nv 1 0 0 # trailing\nv 0 1 0\nvt 0 0\nvn 0 0 1\n"
            "f 1/1/1 2//1 3\nf -3 -2 \\\n -1\nf 4 2 3\nv 1 1 0\n")
    p, f = pkg.io.parse_obj(text)
    assert p.shape == (4, 3) and f.tolist() == [[0, 1, 2], [0, 1, 2], [3, 1, 2]]
    for bad, kind in (("v 0 0\n", "Parse"), ("v 0 0 0\nf 1 1\n", "Parse"), ("v 0 0 0\nf 0 1 1\n", "Parse"),
                      ("v 0 0 0\nf 1 1 5\n", "Parse"), ("v 0 0 nan\n", "Parse"),
                      ("v 0 0 0\nv 1 0 0\nv 0 1 0\nv 1 1 0\nf 1 2 4 3\n", "PolygonFace")):
        with pytest.raises(pkg.RsError) as e:
            pkg.io.parse_obj(bad)
        assert e.value.kind == kind, bad
    with pytest.raises(pkg.RsError) as e:
        pkg.io.parse_obj("v 0 0 0\nv 0 0 0\nf 1 2 -3\n")
    assert "line 3" in e.value.detail


@pytest.mark.parametrize("name", fx.AUTHORS_NAMES)
def test_authors_obj_round_trip(name, tmp_path):
    pkg = load_package()
    p, f = pkg.io.read_obj(f"{fx.AUTHORS_DIR}/{name}.obj")
    p0, f0 = fx.read_obj(f"{fx.AUTHORS_DIR}/{name}.obj")
    assert np.array_equal(p, p0) and np.array_equal(f, f0)
    path = tmp_path / "m.obj"
    pkg.io.write_obj(path, p, f)
    p2, f2 = pkg.io.read_obj(path)
    assert np.array_equal(p2, p) and np.array_equal(f2, f)
    with pytest.raises(pkg.RsError) as e:
        pkg.io.read_obj(tmp_path / "missing.obj")
    assert e.value.kind == "Io"


def test_session_options():
    pkg = load_package()
    o, step = pkg.api._options({"alpha": 4.0, "beta": 8.0, "chi": 0.25, "gmres_tolerance": 1e-8, "start": "absolute",
                                "start_tau": 0.5, "max_halvings": 3, "energy_eval": "exact", "low_order": False})
    assert step.exponents == pkg.Exponents(4.0, 8.0) and step.energy_exact and step.saddle.chi == 0.25
    assert step.saddle.c_lo == 0.0 and step.gmres.tolerance == 1e-8 and step.max_halvings == 3
    assert step.start == pkg.StartRule.absolute(0.5)
    o, step = pkg.api._options({"constraints": [{"type": "barycenter"}, {"type": "area"},
                                                {"type": "volume", "target": 2.5}], "newton_steps": 1})
    assert step.constraints == (("area", None), ("volume", 2.5)) and step.newton == ("fixed", 1)
    assert pkg.api._options({})[1].constraints == () and pkg.api._options({})[1].newton == ("tol", 5)
    # the CCD cap is on by default, as in the reference (repulsive_surfaces.rs:1048)
    assert pkg.api._options({})[1].ccd == pkg.CcdOptions()
    assert pkg.api._options({"ccd": False})[1].ccd is None
    c = pkg.api._options({"ccd_gap": 0.1, "ccd_max_iterations": 7, "ccd_close_factor": 0.25})[1].ccd
    assert (c.gap, c.max_iterations, c.close_factor, c.reference_length) == (0.1, 7, 0.25, None)
    # remeshing is on by default with the reference's RemeshOptions (repulsive_surfaces.rs:1062-1111)
    assert pkg.api._options({})[0]["_remesh"] == pkg.RemeshOptions()
    assert pkg.api._options({"remesh": False})[0]["_remesh"] is None
    r = pkg.api._options({"remesh_split": 1.8, "remesh_collapse": 0.4, "remesh_rounds": 2, "remesh_rho": 0.25,
                          "remesh_flip_cap": 7, "remesh_order": "smooth_then_flip",
                          "remesh_ops": ["flip", "smooth"]})[0]["_remesh"]
    assert (r.split, r.collapse, r.rounds, r.rho, r.flip_cap_per_vertex, r.order) == (1.8, 0.4, 2, 0.25, 7,
                                                                                       "smooth_then_flip")
    assert r.enabled_ops == pkg.EnabledOps(False, False, True, True)
    for bad, kind in (({"remesh_order": "sideways"}, "InvalidOption"), ({"remesh_ops": ["twist"]}, "InvalidOption"),
                      ({"remesh_ops": "flip"}, "InvalidOption"), ({"device": "cuda"}, "Unsupported"),
                      ({"metric": "h1"}, "Unsupported"), ({"penalties": [1]}, "Unsupported"),
                      ({"constraints": [{"type": "area"}]}, "Unsupported"),
                      ({"constraints": [{"type": "barycenter"}, {"type": "pins", "vertices": [0]}]}, "Unsupported"),
                      ({"constraints": [{"type": "barycenter"}, {"type": "area", "component": 0}]}, "Unsupported"),
                      ({"constraints": [{"type": "barycenter"}, {"type": "area", "factor": 1.1, "steps": 3}]},
                       "Unsupported"),
                      ({"constraints": [{"type": "normal"}]}, "InvalidOption"), ({"barycenter": False}, "Unsupported"),
                      ({"no_such_option": 1}, "InvalidOption"), ({"start": "absolute"}, "InvalidOption")):
        with pytest.raises(pkg.RsError) as e:
            pkg.api._options(bad)
        assert e.value.kind == kind, bad
