"""Objects imported from OBJ files with MTL materials (create_object,
src/xtcore/parseutil.cc:588-682, over tinyobjloader 1.0) and the procedural
meshes gen(icosahedron) / gen(plane) (parseutil.cc:295-314).

Where the compiled reference is built, the loader's objects, the traces and
the deterministic images are compared with it bit for bit, and the
pathtracer statistically. The reference's gen(icosahedron) reads
uninitialised normal and texcoord indices (icosahedron.cc:62-67) and
crashes; its scenes are compared with the icosahedron swapped for a
gen(plane), and the icosahedron itself is checked against its closed form."""
import json
import os
import subprocess
import sys

import numpy as np
import pytest

import xtracer_amd as xa
from oracle import oracle as O
from tests.helpers import external_scene

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
needs_ref = pytest.mark.skipif(not O.ref_available(), reason="reference checker (oracle/_ref) not built")


def _no_ico(path):
    text = open(path).read().replace("source = gen(icosahedron)", "source = gen(plane)")
    with open(path, "w") as f:
        f.write(text)
    return path


def test_shapes_materials_and_names(tmp_path):
    """Shapes split at o / g; a shape takes its first face's material; a
    group whose faces an earlier usemtl moved out is dropped (tinyobj 1.0);
    names are k + prefix + shape name; a repeated MTL name keeps its first
    material; emissive and textured MTL channels."""
    s = xa.Scene(external_scene(str(tmp_path / "e.scn")))
    objs = {o["name"]: o for o in s.objects()}
    ext = [n for n in objs if "ext_" in n]
    assert ext == ["1ext_floor", "2ext_lamps", "3ext_mixed", "4ext_tail"]
    assert [objs[n]["prim_count"] for n in ext] == [8, 2, 6, 5]
    assert objs["grid"]["prim_count"] == 18 and objs["ico"]["prim_count"] == 20
    mats = s.arrays()["materials"]
    assert len(mats) == 6  # grey + red, shiny, lamp, wood, glow: the second "red" is not added
    m = {n: mats[objs[n]["material"]] for n in ext}
    np.testing.assert_array_equal(m["1ext_floor"]["kd"], np.float32([0.8, 0.1, 0.1]))  # the first "red"
    assert m["1ext_floor"]["type"] == 2 and m["1ext_floor"]["exponent"] == np.float32(12)
    assert m["2ext_lamps"]["type"] == 3
    np.testing.assert_array_equal(m["2ext_lamps"]["ke"], np.float32([3, 2.5, 2]))
    assert m["3ext_mixed"]["type"] == 3  # glow, its first face's material
    assert m["4ext_tail"]["type"] == 2  # wood
    assert ((mats["type"] == 2) & (mats["exponent"] == np.float32(40))).sum() == 1  # shiny
    smp = s.samplers()
    wood, glow = smp["mat_smp"][objs["4ext_tail"]["material"]], smp["mat_smp"][objs["3ext_mixed"]["material"]]
    assert wood[0] != 0xffffffff and wood[1] == 0xffffffff and glow[2] != 0xffffffff
    tex = smp["tex_info"][smp["samplers"][glow[2]][2]]
    o, w, h = int(tex[0]), int(tex[1]), int(tex[2])
    got = smp["texels"][o:o + w * h].reshape(h, w, 4)
    # MTL textures are mirrored left to right (create_sampler, parseutil.cc:577)
    assert np.array_equal(got, xa.read_image(tmp_path / "maps" / "glow.tga")[:, ::-1])


def test_gen_icosahedron_closed_form(tmp_path):
    s = xa.Scene(external_scene(str(tmp_path / "e.scn")))
    k = [i for i, o in enumerate(s.objects()) if o["name"] == "ico"][0]
    o = s.objects()[k]
    tris = s.arrays()["prims64"][o["first_prim"]:o["first_prim"] + o["prim_count"]].reshape(-1, 3, 3)
    v = (tris - np.array([-1.4, 1.2, -1.0])) / 0.4  # undo the modifiers
    np.testing.assert_allclose(np.linalg.norm(v, axis=2), 1.0, rtol=1e-6)
    edges = np.linalg.norm(v - np.roll(v, 1, axis=1), axis=2)
    np.testing.assert_allclose(edges, 1.0514622, rtol=1e-5)  # edge of the unit-radius icosahedron
    assert len(np.unique(np.round(v.reshape(-1, 3), 5), axis=0)) == 12


def test_missing_obj_adds_nothing(tmp_path):
    path = external_scene(str(tmp_path / "e.scn"))
    os.remove(tmp_path / "model.obj")
    s = xa.Scene(path)
    assert not [o for o in s.objects() if "ext_" in o["name"]]


def test_missing_mtl_texture_fails(tmp_path):
    path = external_scene(str(tmp_path / "e.scn"))
    os.remove(tmp_path / "maps" / "wood.png")
    with pytest.raises(xa.XtError):
        xa.Scene(path)


_CHILD = r"""
import json, sys
import numpy as np
sys.path.insert(0, {root!r})
import xtracer_amd as xa
from oracle import oracle as O
from tests.helpers import camera_dict_of
path = {path!r}
s = xa.Scene(path)
r = O.RefScene(path)
po, ro = s.objects(), r.objects()
res = {{"names": [p["name"] for p in po] == [q["name"] for q in ro]}}
orc = O.OracleScene(s)
rng = np.random.default_rng(5)
o = rng.uniform(-4, 4, (40000, 3))
d = rng.normal(size=(40000, 3)); d /= np.linalg.norm(d, axis=1)[:, None]
t1, ob1, _, p1, n1 = orc.trace(o, d)
t2, ob2, p2, n2 = r.trace(o, d)
res["trace"] = bool(np.array_equal(t1, t2) and np.array_equal(ob1, ob2) and np.array_equal(n1[ob1 >= 0], n2[ob2 >= 0]))
res["hits"] = int((ob1 >= 0).sum())
bad = {{}}
for integ in (0, 1, 4, 5, 6):
    ref = r.render(integ, 64, 48, samples=4, threads=1)
    mine, _ = orc.render(integ, 64, 48, camera_dict_of("pinhole"), samples=4)
    bad[integ] = [int((ref.view(np.uint32) != mine.view(np.uint32)).any(axis=2).sum()),
                  int((ref[..., :3] > 0).any(axis=2).sum())]
res["bad"] = bad
spp = 256
ref = r.render(3, 24, 16, samples=spp, rdepth=6, threads=1, seed=7)
_, _, mean, m2 = orc.render(3, 24, 16, camera_dict_of("pinhole"), samples=spp, rdepth=6, seed=3, stats=True)
se = np.sqrt(2 * (m2 / (spp - 1)) / spp)
res["inside"] = float((np.abs(ref[..., :3] - mean) <= 2.576 * se + 1e-5).mean())
print(json.dumps(res))
"""


@needs_ref
def test_external_scene_matches_reference(tmp_path):
    """Object names and scan order, closest hits (t, object, normal),
    stencil / depth / normal / uv / emission images bit for bit (the
    emission image reads the MTL texture of an emissive shape), and the
    pathtracer within its 99% Monte Carlo bound."""
    path = _no_ico(external_scene(str(tmp_path / "e.scn")))
    code = _CHILD.format(root=ROOT, path=path)
    out = json.loads(subprocess.check_output([sys.executable, "-c", code], timeout=600).decode().splitlines()[-1])
    assert out["names"] and out["trace"] and out["hits"] > 10000, out
    for integ, (bad, lit) in out["bad"].items():
        assert bad == 0, (integ, out["bad"])
        assert lit > 100, (integ, out["bad"])
    assert out["inside"] >= 0.97, out


def test_extrude_radii(tmp_path):
    """extrude (extrude.cc:9-27) puts every vertex at 10 * (red channel of
    the cubemap in its direction) from the origin."""
    from tests.helpers import extrude_scene
    s = xa.Scene(extrude_scene(str(tmp_path / "x.scn")))
    v = s.arrays()["prims64"][:, :9].reshape(-1, 3)
    r = np.linalg.norm(v, axis=1)
    assert r.max() <= 10.0 + 1e-5 and len(np.unique(np.round(r, 4))) > 5


_EXTRUDE = r"""
import json, sys
import numpy as np
sys.path.insert(0, {root!r})
import xtracer_amd as xa
from oracle import oracle as O
from xtracer_amd.scenes import camera_dict
path = {path!r}
s = xa.Scene(path)
r = O.RefScene(path)
orc = O.OracleScene(s)
rng = np.random.default_rng(5)
o = rng.uniform(-6, 6, (40000, 3))
d = rng.normal(size=(40000, 3)); d /= np.linalg.norm(d, axis=1)[:, None]
t1, ob1, _, p1, n1 = orc.trace(o, d)
t2, ob2, p2, n2 = r.trace(o, d)
res = {{"trace": bool(np.array_equal(t1, t2) and np.array_equal(ob1, ob2)), "hits": int((ob1 >= 0).sum())}}
cam = camera_dict(position=(0.5, 2.0, -9.0), target=(0.0, 0.3, 0.0), fov=50.0)
res["bad"] = [int((r.render(k, 64, 48, samples=4, threads=1).view(np.uint32)
                   != orc.render(k, 64, 48, cam, samples=4)[0].view(np.uint32)).any(axis=2).sum()) for k in (0, 1, 4)]
print(json.dumps(res))
"""


@needs_ref
def test_extrude_matches_reference(tmp_path):
    """A rotated, scaled mesh extruded by a cubemap: hits and stencil /
    depth / normal images equal the compiled reference's bit for bit."""
    from tests.helpers import extrude_scene
    path = extrude_scene(str(tmp_path / "x.scn"))
    code = _EXTRUDE.format(root=ROOT, path=path)
    out = json.loads(subprocess.check_output([sys.executable, "-c", code], timeout=600).decode().splitlines()[-1])
    assert out["trace"] and out["hits"] > 10000 and out["bad"] == [0, 0, 0], out
