"""The C-ABI library loads and exports every entry point include/xtracer_amd.h
declares; host-side calls (no GPU) behave like the reference loader."""
import ctypes
import os
import re

import numpy as np
import pytest

import xtracer_amd as xa
from xtracer_amd import _native, scenes

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    text = open(os.path.join(ROOT, "include", "xtracer_amd.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(xt_[a-z0-9_]+)\s*\(", text)))


def test_exports_every_declared_symbol():
    lib = ctypes.CDLL(_native.LIB_PATH)
    names = declared_functions()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert sorted(_native.SYMBOLS) == names


def test_missing_scene_is_io_error(tmp_path):
    with pytest.raises(xa.XtError) as e:
        xa.Scene(str(tmp_path / "nope.scn"))
    assert e.value.code == 2


def test_parse_error_reports_line(tmp_path):
    p = tmp_path / "bad.scn"
    p.write_text("title = x\ncamera = {\n\tdefault = {\n\t\ttype thin-lens\n\t}\n}\n")
    with pytest.raises(xa.XtError) as e:
        xa.Scene(str(p))
    assert e.value.code == 2 and "line 4" in str(e.value) and "syntax" in str(e.value)


def test_unbalanced_groups(tmp_path):
    p = tmp_path / "bad.scn"
    p.write_text("camera = {\n\tdefault = {\n\t\ttype = thin-lens\n\t}\n")
    with pytest.raises(xa.XtError) as e:
        xa.Scene(str(p))
    assert "unbalanced" in str(e.value)


def test_unknown_material_fails_load(tmp_path):
    # parseutil.cc:436-447 returns 0 for an unknown type, create_material then fails the load
    p = tmp_path / "m.scn"
    p.write_text("material = {\n\tm = {\n\t\ttype = velvet\n\t}\n}\n")
    with pytest.raises(xa.XtError) as e:
        xa.Scene(str(p))
    assert e.value.code == 3


def test_host_arrays_and_object_order():
    s = xa.Scene(scenes.spheres(seed=0))
    info = s.info()
    objs = s.objects()
    assert info.n_objects == 66 and info.n_planes == 1 and info.n_prims == 65
    # objects are sorted by first-seen name order: camera, geometry, material
    # and object sections in that order, each alphabetically
    assert [o["handle"] for o in objs] == sorted(o["handle"] for o in objs)
    names = [o["name"] for o in objs]
    assert names[:2] == ["floor", "lamp"] and names[2] == "ball000"
    arr = s.arrays()
    lamp = objs[1]
    assert arr["prims"]["tf"][lamp["first_prim"]] & 15 == 1  # a sphere ...
    assert (arr["prims"]["tf"][lamp["first_prim"]] >> 4) & 4  # ... of radius 0: never hit
    assert np.all(arr["prims"]["object"][arr["prims"]["object"] >= 0] < info.n_objects)


def test_camera_basis_matches_closed_form():
    c = xa.make_camera((0, 0, 0), (0, 0, 1), (0, 1, 0), 90.0)
    # up x z = (1,0,0), (1,0,0) x z = (0,-1,0): rows (rx.x, ry.x, rz.x) ...
    assert list(c.basis) == [1.0, 0.0, 0.0, 0.0, -1.0, 0.0, 0.0, 0.0, 1.0]
    assert abs(c.dir_z - 1.0) < 1e-12 and c.thin_lens == 0


def test_symbol_expansion_and_mods(tmp_path):
    p = tmp_path / "s.scn"
    p.write_text("r = 2.5\ngeometry = {\n\tb = {\n\t\ttype = sphere\n\t\tradius = <r>\n"
                 "\t\tposition = vec3(1,2,3)\n\t}\n}\nobject = {\n\to = {\n\t\tgeometry = b\n\t\tmaterial = m\n\t}\n}\n"
                 "material = {\n\tm = {\n\t\ttype = lambert\n\t}\n}\n")
    s = xa.Scene(str(p))
    assert s.object_params(0)[3] == 2.5
    s2 = xa.Scene(str(p), mods=["geometry.b.radius:0.75"])
    assert s2.object_params(0)[3] == 0.75


def test_include(tmp_path):
    (tmp_path / "geo.inc").write_text("\tb = {\n\t\ttype = sphere\n\t\tradius = 1\n\t}\n")
    p = tmp_path / "s.scn"
    p.write_text("geometry = {\n%include geo.inc\n}\nobject = {\n\to = {\n\t\tgeometry = b\n\t\tmaterial = m\n\t}\n}\n"
                 "material = {\n\tm = {\n\t\ttype = lambert\n\t}\n}\n")
    s = xa.Scene(str(p))
    assert s.objects()[0]["geom_type"] == 1 and s.object_params(0)[3] == 1.0


def test_mesh_face_indices_out_of_range(tmp_path):
    """A mesh OBJ whose faces index past the vertex, texcoord or normal
    arrays loads and builds: such an index reads as absent (vertex (0, 0, 0),
    default texcoord, no vertex normal), as for imported objects."""
    (tmp_path / "m.obj").write_text("v 0 0 0\nv 1 0 0\nv 0 1 0\nvt 0.5 0.5\nvn 0 0 1\n"
                                    "f 1 2 3\nf 1 2 9999\nf 1/7/9 2/1/1 3//-5\nf -9 2 3\n")
    (tmp_path / "m.scn").write_text(
        "geometry = {\n\tg = {\n\t\ttype = mesh\n\t\tsource = m.obj\n\t}\n}\n"
        "object = {\n\to = {\n\t\tgeometry = g\n\t\tmaterial = m\n\t}\n}\n"
        "material = {\n" + scenes._material("m", "lambert", kd=(0.5, 0.5, 0.5)) + "\n}\n")
    s = xa.Scene(str(tmp_path / "m.scn"))
    arr = s.arrays()
    assert len(arr["prims"]) == 4
    order = np.argsort(arr["prim_local"])
    v = arr["prims64"][order].reshape(4, 3, 3)
    assert np.array_equal(v[0], [[0, 0, 0], [1, 0, 0], [0, 1, 0]])
    assert np.array_equal(v[1][2], [0, 0, 0])
    assert np.array_equal(v[3][0], [0, 0, 0])
    tc, _ = s.texcoords()
    assert np.array_equal(tc[order][2], [0, 0, 0.5, 0.5, 0, 0])
