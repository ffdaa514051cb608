"""Shared helpers for the parity tests."""
import ctypes

import numpy as np

import xtracer_amd as xa
from xtracer_amd import scenes


def soup_to_object_local(scene_arrays, prim):
    """Maps device hit ids to (object index, index inside the geometry)."""
    prims = scene_arrays["prims"]
    planes = scene_arrays["planes"]
    local = scene_arrays["prim_local"]
    prim = np.asarray(prim, np.uint64)
    obj = np.full(len(prim), -1, np.int64)
    loc = np.full(len(prim), -1, np.int64)
    soup = prim < len(prims)
    obj[soup] = prims["object"][prim[soup].astype(np.int64)]
    loc[soup] = local[prim[soup].astype(np.int64)]
    pl = (prim >= xa.HIT_PLANE_BASE) & (prim != xa.HIT_NONE)
    if pl.any():
        obj[pl] = planes["object"][(prim[pl] - xa.HIT_PLANE_BASE).astype(np.int64)]
        loc[pl] = 0
    return obj, loc


def camera_rays(cam, width, height):
    """Pinhole primary rays (pixel centres) in FP64, as Perspective::get_primary_ray."""
    pos = np.array(cam.position)
    m = np.array(cam.basis).reshape(3, 3)
    ys, xs = np.mgrid[0:height, 0:width]
    x = (xs + 0.5).astype(np.float32).astype(np.float64).ravel()
    y = (ys + 0.5).astype(np.float32).astype(np.float64).ravel()
    w, h = float(np.float32(width)), float(np.float32(height))
    aspect = w / h
    d = np.stack([(2.0 * x / w) - 1.0, ((2.0 * y / h) - 1.0) / aspect, np.full(x.shape, cam.dir_z)], 1)
    d = d @ m.T
    d /= np.linalg.norm(d, axis=1)[:, None]
    o = np.broadcast_to(pos, d.shape).copy()
    return o, d


def mixed_rays(scene_path, n, seed=0, box=(-3, 6)):
    """Random rays: half from random points in a box in random directions, half
    camera-like rays through the scene, as FP32 (origins, unit directions)."""
    rng = np.random.default_rng(seed)
    o = rng.uniform(box[0], box[1], (n, 3))
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1)[:, None]
    o = o.astype(np.float32)
    d = d.astype(np.float32)
    d = (d / np.linalg.norm(d.astype(np.float64), axis=1)[:, None]).astype(np.float32)
    return o, d


def scene_path(kind, **kw):
    return scenes.SCENES[kind][0](**kw)


def glow_scene(path):
    """A pinhole scene with emissive and non-emissive materials in view: an
    emissive sphere, a Lambert sphere that also carries an emissive sampler,
    a plain Lambert sphere, a Phong triangle and a Lambert floor plane."""
    from tests import scn
    scn.write(path,
              geometry={"floor": dict(type="plane", normal=(0, 1, 0), distance=0.0),
                        "a": dict(type="sphere", position=(-1.2, 0.6, 0.0), radius=0.6),
                        "b": dict(type="sphere", position=(0.2, 0.5, 0.4), radius=0.5),
                        "c": dict(type="sphere", position=(1.4, 0.4, -0.3), radius=0.4),
                        "tri": dict(type="triangle", v=[(-2.0, 0.0, 1.5), (2.0, 0.0, 1.5), (0.0, 2.2, 1.8)])},
              objects={"floor": ("floor", "ground"), "a": ("a", "lamp"), "b": ("b", "glowing"),
                       "c": ("c", "matte"), "tri": ("tri", "shiny")},
              materials={"ground": dict(type="lambert", kd=(0.5, 0.5, 0.5)),
                         "lamp": dict(type="emissive", ke=(4.0, 3.5, 2.0)),
                         "glowing": dict(type="lambert", kd=(0.2, 0.6, 0.2), ke=(0.3, 1.25, 0.5)),
                         "matte": dict(type="lambert", kd=(0.7, 0.2, 0.2)),
                         "shiny": dict(type="phong", kd=(0.3, 0.3, 0.6), ks=(0.5, 0.5, 0.5), reflectance=0.5,
                                       exponent=20.0)},
              camera=dict(position=(0.3, 1.4, -4.5), target=(0.0, 0.6, 0.5), fov=50.0),
              env=("gradient", (1.0, 1.0, 1.0), (0.5, 0.7, 1.0)))
    return path


CAMERAS = {
    "pinhole": dict(position=(0.3, 1.4, -6.5), target=(0.0, 0.6, 0.5), fov=50.0),
    "erp": dict(type="erp", position=(0.1, 0.9, -1.6), orientation=(0.3, -0.7, 1.1)),
    "ods": dict(type="ods", position=(0.1, 0.9, -1.6), orientation=(-0.2, 0.4, 0.05), ipd=0.25),
    "cubemap": dict(type="cubemap", position=(0.1, 0.9, -1.6)),
}


def camera_dict_of(kind):
    """CAMERAS[kind] in the form the oracle takes, float32-rounded like the loader."""
    from xtracer_amd.scenes import camera_dict
    if kind == "pinhole":
        return camera_dict(**CAMERAS[kind])
    cam = dict(CAMERAS[kind])
    for k in ("position", "orientation"):
        if k in cam:
            cam[k] = [float(np.float32(c)) for c in cam[k]]
    if "ipd" in cam:
        cam["ipd"] = float(np.float32(cam["ipd"]))
    return cam


def camera_scene(path, kind):
    """The glow scene's objects and materials, plus a dielectric sphere, seen
    by camera CAMERAS[kind]: a pinhole, or the all-around erp, ods or cubemap
    from inside the scene."""
    from tests import scn
    scn.write(path,
              geometry={"floor": dict(type="plane", normal=(0, 1, 0), distance=0.0),
                        "a": dict(type="sphere", position=(-1.2, 0.6, 0.0), radius=0.6),
                        "b": dict(type="sphere", position=(0.2, 0.5, 0.4), radius=0.5),
                        "c": dict(type="sphere", position=(1.4, 0.4, -0.3), radius=0.4),
                        "glass": dict(type="sphere", position=(-0.4, 0.8, -3.0), radius=0.7),
                        "tri": dict(type="triangle", v=[(-2.0, 0.0, 1.5), (2.0, 0.0, 1.5), (0.0, 2.2, 1.8)])},
              objects={"floor": ("floor", "ground"), "a": ("a", "lamp"), "b": ("b", "glowing"),
                       "c": ("c", "matte"), "glass": ("glass", "glass"), "tri": ("tri", "shiny")},
              materials={"ground": dict(type="lambert", kd=(0.5, 0.5, 0.5)),
                         "lamp": dict(type="emissive", ke=(4.0, 3.5, 2.0)),
                         "glowing": dict(type="lambert", kd=(0.2, 0.6, 0.2), ke=(0.3, 1.25, 0.5)),
                         "matte": dict(type="lambert", kd=(0.7, 0.2, 0.2)),
                         "glass": dict(type="dielectric", ior=1.5),
                         "shiny": dict(type="phong", kd=(0.3, 0.3, 0.6), ks=(0.5, 0.5, 0.5), reflectance=0.5,
                                       exponent=20.0)},
              camera=CAMERAS[kind],
              env=("gradient", (1.0, 1.0, 1.0), (0.5, 0.7, 1.0)))
    return path


def uv_scene(path, camera="pinhole"):
    """Texture-coordinate sources in view: spheres with and without u/v
    scales (one seen from inside by the all-around cameras), a scaled floor
    plane, a triangle with corner texcoords, and a mesh whose faces have vt
    indices, or none (Triangle() defaults)."""
    import os
    from tests import scn
    d = os.path.dirname(str(path))
    rng = np.random.default_rng(17)
    tris = []
    for k in range(12):
        c = np.array([-2.5 + 0.45 * k, 0.4 + 0.1 * (k % 3), 2.2])
        tris.append([c + [-0.2, -0.2, 0], c + [0.2, -0.2, 0.05], c + [0.0, 0.25, -0.05]])
    uvs = [None if k % 4 == 0 else [tuple(rng.uniform(-1.5, 2.5, 2)) for _ in range(3)] for k in range(12)]
    obj = scn.write_obj_uv(os.path.join(d, "uvmesh.obj"), tris, uvs)
    scn.write(path,
              geometry={"floor": dict(type="plane", normal=(0, 1, 0), distance=0.0, u_scale=0.7, v_scale=1.3),
                        "a": dict(type="sphere", position=(-1.2, 0.6, 0.0), radius=0.6, u_scale=1.5, v_scale=0.8),
                        "b": dict(type="sphere", position=(0.2, 0.5, 0.4), radius=0.5),
                        "tri": dict(type="triangle", v=[(-2.0, 0.0, 1.5), (2.0, 0.0, 1.5), (0.0, 2.2, 1.8)],
                                    tc=[(0.1, 0.9), (0.8, 0.2), (0.45, -0.3)]),
                        "m": dict(type="mesh", source=os.path.basename(obj))},
              objects={"floor": ("floor", "ground"), "a": ("a", "ground"), "b": ("b", "ground"),
                       "tri": ("tri", "ground"), "m": ("m", "ground")},
              materials={"ground": dict(type="lambert", kd=(0.5, 0.5, 0.5))},
              camera=CAMERAS[camera],
              env=("gradient", (1.0, 1.0, 1.0), (0.5, 0.7, 1.0)))
    return path


def texture_files(d):
    """Small texture files of each decoded format in directory d (odd sizes,
    so bilinear wraps and nearest clamps are exercised)."""
    import os
    from tests import imgfiles as I
    rng = np.random.default_rng(23)
    yy, xx = np.mgrid[0:9, 0:13]
    rgb = np.stack([(xx * 19) % 256, (yy * 27) % 256, ((xx + yy) % 2) * 200 + 30], -1).astype(np.uint8)
    rgba = np.concatenate([rng.integers(0, 256, (7, 10, 3)), rng.integers(0, 256, (7, 10, 1))], -1).astype(np.uint8)
    f = {"rgb_png": I.png(os.path.join(d, "tex_rgb.png"), rgb),
         "rgba_png": I.png(os.path.join(d, "tex_rgba.png"), rgba, filters=(4, 3, 1)),
         "tga": I.tga(os.path.join(d, "tex.tga"), rng.integers(0, 256, (6, 11, 3)).astype(np.uint8), rle=True),
         "bmp": I.bmp(os.path.join(d, "tex.bmp"), rng.integers(0, 256, (5, 6, 3)).astype(np.uint8)),
         "erp": I.png(os.path.join(d, "erp.png"), rng.integers(0, 256, (16, 32, 3)).astype(np.uint8))}
    for k, face in enumerate(["posx", "negx", "posy", "negy", "posz", "negz"]):
        img = np.zeros((8, 8, 3), np.uint8)
        img[..., k % 3] = 60 + 30 * k
        img[..., (k + 1) % 3] = (np.arange(8)[None, :] * 31) % 256
        img[..., (k + 2) % 3] = (np.arange(8)[:, None] * 29) % 256
        f[face] = I.ppm(os.path.join(d, f"{face}.ppm"), img) if k % 2 else I.tga(os.path.join(d, f"{face}.tga"), img)
    return {k: os.path.basename(v) for k, v in f.items()}


def texture_scene(path, camera="pinhole", variant="emit", env="erp"):
    """uv_scene's geometry with textured materials. variant "emit": every
    object emissive through a texture (nearest, bilinear, flips; texcoords
    outside [0, 1] on the mesh), gradient or ERP sampler read at its texcoord; variant "lit": textured
    lambert / phong / blinn-phong diffuse and specular under one textured
    light; variant "env": no objects, the environment alone. env: "erp",
    "cubemap" or "gradient"."""
    import os
    from tests import scn
    d = os.path.dirname(str(path))
    uv_scene(path, camera)  # writes the mesh file
    f = texture_files(d)
    cube = {k: f[k] for k in ("posx", "negx", "posy", "negy", "posz", "negz")}
    tex = lambda key, **kw: dict(type="texture", source=f[key], **kw)
    if variant == "emit":
        mats = {"m_floor": dict(type="emissive", ke=tex("rgb_png", filtering="bilinear")),
                "m_a": dict(type="emissive", ke=tex("tga", flip_x=True)),
                "m_b": dict(type="emissive", ke=tex("rgba_png", filtering="bilinear", flip_y=True)),
                "m_tri": dict(type="emissive", ke=dict(type="erp", source=f["erp"])),
                "m_m": dict(type="emissive", ke=tex("bmp")),
                "m_t2": dict(type="emissive", ke=dict(type="gradient", a=(0.9, 0.2, 0.1), b=(0.1, 0.3, 1.0)))}
    else:
        mats = {"m_floor": dict(type="lambert", kd=tex("rgb_png", filtering="bilinear")),
                "m_a": dict(type="phong", kd=tex("bmp"), ks=tex("tga", flip_y=True), exponent=20.0),
                "m_b": dict(type="emissive", ke=tex("rgba_png", flip_x=True)),
                "m_tri": dict(type="blinn_phong", kd=tex("rgba_png", filtering="bilinear"), ks=(0.3, 0.3, 0.3),
                              exponent=8.0),
                "m_m": dict(type="lambert", kd=dict(type="cubemap", **cube)),
                "m_t2": dict(type="lambert", kd=dict(type="gradient", a=(0.9, 0.2, 0.1), b=(0.1, 0.3, 1.0)))}
    envd = {"erp": dict(type="erp", source=f["erp"]), "cubemap": dict(type="cubemap", **cube),
            "gradient": ("gradient", (1.0, 1.0, 1.0), (0.5, 0.7, 1.0))}[env]
    geo = {"floor": dict(type="plane", normal=(0, 1, 0), distance=0.0, u_scale=0.7, v_scale=1.3),
           "a": dict(type="sphere", position=(-1.2, 0.6, 0.0), radius=0.6, u_scale=1.5, v_scale=0.8),
           "b": dict(type="sphere", position=(0.2, 0.5, 0.4), radius=0.5),
           "tri": dict(type="triangle", v=[(-2.0, 0.0, 1.5), (2.0, 0.0, 1.5), (0.0, 2.2, 1.8)],
                       tc=[(0.1, 0.9), (0.8, 0.2), (0.45, -0.3)]),
           "m": dict(type="mesh", source="uvmesh.obj"),
           "t2": dict(type="triangle", v=[(1.0, 0.05, -1.0), (2.6, 0.05, -0.6), (1.8, 1.2, -0.8)],
                      tc=[(-0.7, 1.6), (2.3, -0.4), (1.2, 2.7)])}
    objs = {"floor": ("floor", "m_floor"), "a": ("a", "m_a"), "b": ("b", "m_b"), "tri": ("tri", "m_tri"),
            "m": ("m", "m_m"), "t2": ("t2", "m_t2")}
    if variant == "env":
        geo, objs, mats = {}, {}, {}
    scn.write(path, geometry=geo, objects=objs, materials=mats, camera=CAMERAS[camera], env=envd)
    return path


def external_scene(path, camera="pinhole"):
    """An object imported from an OBJ file with MTL materials (prefix
    "ext_"): shapes split by `o` / `g`, `usemtl` switching inside a group
    (the shape takes its first face's material; a group whose earlier faces
    tinyobj drops), quads, vt / vn indices,
    colour, emissive, textured (map_Kd, map_Ke, a backslash path) and
    repeated material names; beside it a procedural icosahedron (scaled and
    moved) and a gen(plane) of resolution 3."""
    import os
    from tests import imgfiles as I
    from tests import scn
    d = os.path.dirname(str(path))
    os.makedirs(os.path.join(d, "maps"), exist_ok=True)
    rng = np.random.default_rng(29)
    I.png(os.path.join(d, "maps", "wood.png"), rng.integers(0, 256, (6, 9, 3)).astype(np.uint8))
    I.tga(os.path.join(d, "maps", "glow.tga"), rng.integers(0, 256, (5, 4, 3)).astype(np.uint8))
    with open(os.path.join(d, "model.mtl"), "w") as f:
        f.write("# materials\nnewmtl red\nKd 0.8 0.1 0.1\nKs 0.2 0.2 0.2\nNs 12\n\n"
                "newmtl shiny\nKd 0.2 0.3 0.6\nKs 0.5 0.5 0.4\nNs 40\nNi 1.5\n"
                "newmtl lamp\nKd 0.1 0.1 0.1\nKe 3 2.5 2\n"
                "newmtl wood\nKd 1 1 1\nmap_Kd maps\\wood.png\nKs 0.1 0.1 0.1\n"
                "newmtl glow\nmap_Ke maps/glow.tga\n"
                "newmtl red\nKd 0 1 0\n")
    lines = ["mtllib model.mtl"]
    k = 0
    for gx in range(4):
        for gz in range(3):
            x0, z0 = -2.0 + gx * 1.1, -0.5 + gz * 1.0
            y = 0.3 * ((gx + gz) % 3)
            lines += [f"v {x0:.3f} {y:.3f} {z0:.3f}", f"v {x0 + 0.9:.3f} {y + 0.1:.3f} {z0:.3f}",
                      f"v {x0 + 0.9:.3f} {y + 0.5:.3f} {z0 + 0.8:.3f}", f"v {x0:.3f} {y + 0.4:.3f} {z0 + 0.8:.3f}"]
            k += 4
    lines += ["vt 0 0", "vt 1 0", "vt 1 1", "vt 0 1", "vt 0.25 0.75", "vt 2.5 -1.5", "vn 0 1 0", "vn 0.1 0.9 0.1"]
    quads = [[1 + 4 * q + c for c in range(4)] for q in range(12)]
    def face(q, uv=True, n=False):
        if uv and n:
            return "f " + " ".join(f"{v}/{1 + c % 6}/{1 + c % 2}" for c, v in enumerate(q))
        if uv:
            return "f " + " ".join(f"{v}/{1 + (c + 2) % 6}" for c, v in enumerate(q))
        return "f " + " ".join(str(v) for v in q)
    lines += ["o floor", "usemtl red", face(quads[0]), face(quads[1], n=True),
              "usemtl wood", face(quads[2]), face(quads[3], uv=True, n=True),
              "g lamps", "usemtl lamp", face(quads[4], uv=False),
              "g mixed", "usemtl glow", face(quads[5]), face(quads[6]), "usemtl shiny", face(quads[7]),
              "g dropped", "usemtl red", face(quads[8]), "usemtl shiny",
              "o tail", "usemtl wood", face(quads[9]), "f -4/-1 -3/-2 -2/-3", face(quads[11], n=True)]
    with open(os.path.join(d, "model.obj"), "w") as f:
        f.write("\n".join(lines) + "\n")
    text = open(scn.write(path, geometry={
        "ground": dict(type="plane", normal=(0, 1, 0), distance=-0.2),
        "ball": dict(type="sphere", position=(1.6, 0.5, -1.2), radius=0.4)},
        objects={"ground": ("ground", "grey"), "ball": ("ball", "grey")},
        materials={"grey": dict(type="lambert", kd=(0.5, 0.5, 0.5))},
        camera=CAMERAS[camera], env=("gradient", (1.0, 1.0, 1.0), (0.5, 0.7, 1.0)))).read()
    text = text.replace("geometry = {\n", "geometry = {\n\tico = {\n\t\ttype = mesh\n\t\tsource = gen(icosahedron)\n"
                        "\t\tmodifiers = {\n\t\t\tscale = vec3(0.4,0.4,0.4)\n\t\t\ttranslation = vec3(-1.4,1.2,-1.0)\n"
                        "\t\t}\n\t}\n\tgrid = {\n\t\ttype = mesh\n\t\tsource = gen(plane)\n\t\tresolution = 3\n"
                        "\t\tmodifiers = {\n\t\t\tscale = vec3(1.5,1,1.5)\n\t\t\ttranslation = vec3(0.3,1.6,1.2)\n"
                        "\t\t}\n\t}\n", 1)
    text = text.replace("object = {\n", "object = {\n\tthing = {\n\t\tsource = model.obj\n\t\tprefix = ext_\n\t}\n"
                        "\tico = {\n\t\tgeometry = ico\n\t\tmaterial = grey\n\t}\n"
                        "\tgrid = {\n\t\tgeometry = grid\n\t\tmaterial = grey\n\t}\n", 1)
    with open(path, "w") as f:
        f.write(text)
    return str(path)


def extrude_scene(path, camera="pinhole"):
    """A mesh around the origin (a subdivided octahedron from an OBJ file)
    with the `extrude` modifier: every vertex pushed to 10 * its direction *
    the red channel of a cubemap (texture_files' faces) in that direction;
    moved first by the rotation / scale / translation modifiers."""
    import os
    from tests import scn
    d = os.path.dirname(str(path))
    f = texture_files(d)
    verts = [np.array(v, float) for v in ((1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1))]
    tris = [(0, 2, 4), (2, 1, 4), (1, 3, 4), (3, 0, 4), (2, 0, 5), (1, 2, 5), (3, 1, 5), (0, 3, 5)]
    for _ in range(2):  # two midpoint subdivisions, the new vertices pushed out to the unit sphere
        out = []
        for a, b, c in tris:
            m = []
            for p, q in ((a, b), (b, c), (c, a)):
                v = verts[p] + verts[q]
                verts.append(v / np.linalg.norm(v))
                m.append(len(verts) - 1)
            out += [(a, m[0], m[2]), (m[0], b, m[1]), (m[2], m[1], c), (m[0], m[1], m[2])]
        tris = out
    scn.write_obj(os.path.join(d, "blob.obj"), [[verts[i] for i in t] for t in tris])
    scn.write(path, geometry={"blob": dict(type="mesh", source="blob.obj"),
                              "floor": dict(type="plane", normal=(0, 1, 0), distance=-1.5)},
              objects={"blob": ("blob", "grey"), "floor": ("floor", "grey")},
              materials={"grey": dict(type="lambert", kd=(0.6, 0.6, 0.6))},
              camera=dict(position=(0.5, 2.0, -9.0), target=(0.0, 0.3, 0.0), fov=50.0),
              env=("gradient", (1.0, 1.0, 1.0), (0.5, 0.7, 1.0)))
    faces = "\n".join(f"\t\t\t\t{k} = {f[k]}" for k in ("posx", "negx", "posy", "negy", "posz", "negz"))
    text = open(path).read().replace(
        "\t\tsource = blob.obj\n",
        "\t\tsource = blob.obj\n\t\tmodifiers = {\n\t\t\trotation = vec3(10,20,5)\n\t\t\tscale = vec3(0.9,1.1,1)\n"
        "\t\t\textrude = {\n" + faces + "\n\t\t\t}\n\t\t}\n", 1)
    with open(path, "w") as fh:
        fh.write(text)
    return str(path)


def record_normals(scene, arr, o, d, t, prim, u, v):
    """Normals that the reference's hit record would carry for the device hits
    (prim, t, barycentrics u / v of corners 1 / 2): triangles
    Sum n_i * bc_i with vertex normals (geometric normal when that sum is zero,
    src/xtcore/math/triangle.cc:115-117), spheres (p - c) / r
    (sphere.cc:52), planes their normal as given (plane.cc:41). NaN rows for
    misses."""
    n = len(prim)
    out = np.full((n, 3), np.nan)
    p64 = arr["prims64"]
    prim = np.asarray(prim, np.uint64)
    soup = prim < len(p64)
    idx = prim[soup].astype(np.int64)
    q = p64[idx]
    is_tri = (arr["prims"]["tf"][idx] & 15) == 0
    g = np.cross(q[:, 3:6] - q[:, 0:3], q[:, 6:9] - q[:, 0:3])
    g /= np.maximum(np.linalg.norm(g, axis=1), 1e-300)[:, None]
    res = g.copy()
    vn = scene.vnormals()
    if vn is not None:
        bu, bv = u[soup].astype(np.float64), v[soup].astype(np.float64)
        w = np.stack([1.0 - bu - bv, bu, bv], 1)
        vv = vn[idx].astype(np.float64).reshape(-1, 3, 3)
        s = (vv * w[:, :, None]).sum(1)
        use = (np.abs(vv).sum((1, 2)) > 0) & (np.abs(s).sum(1) > 0)
        res[use] = s[use]
    sph = ~is_tri
    if sph.any():
        pt = o[soup][sph].astype(np.float64) + d[soup][sph].astype(np.float64) * t[soup][sph, None].astype(np.float64)
        res[sph] = (pt - q[sph, 0:3]) / q[sph, 3:4]
    out[np.flatnonzero(soup)] = res
    pl = (prim >= xa.HIT_PLANE_BASE) & (prim != xa.HIT_NONE)
    if pl.any():
        out[pl] = arr["planes"]["normal64"][(prim[pl] - xa.HIT_PLANE_BASE).astype(np.int64)]
    return out


def bary(p, a, b, c):
    v0, v1, v2 = b - a, c - a, p - a
    d00 = (v0 * v0).sum(1)
    d01 = (v0 * v1).sum(1)
    d11 = (v1 * v1).sum(1)
    d20 = (v2 * v0).sum(1)
    d21 = (v2 * v1).sum(1)
    den = d00 * d11 - d01 * d01
    v = (d11 * d20 - d01 * d21) / den
    w = (d00 * d21 - d01 * d20) / den
    return np.stack([1 - v - w, v, w], 1)


# the FP32 test defers every candidate whose smallest barycentric lies within
# 2e-5 of an edge; disagreements with the FP64 reference must lie in that band
EDGE_BAND = 2e-5


def plane_point(arr, prim, o, d):
    """FP64 point where each ray meets the plane of soup triangle `prim`
    (the ray's origin for non-triangles)."""
    out = o.astype(np.float64).copy()
    p64 = arr["prims64"]
    ok = prim < len(p64)
    if ok.any():
        q = p64[prim[ok].astype(np.int64)]
        nrm = np.cross(q[:, 3:6] - q[:, 0:3], q[:, 6:9] - q[:, 0:3])
        oo, dd = o[ok].astype(np.float64), d[ok].astype(np.float64)
        den = (nrm * dd).sum(1)
        tt = ((nrm * (q[:, 0:3] - oo)).sum(1)) / np.where(den != 0, den, 1e-300)
        out[np.flatnonzero(ok)] = oo + dd * tt[:, None]
    return out


def edge_band(arr, prim, point, tol=EDGE_BAND):
    """True where `point` lies within `tol` (barycentric) of an edge of soup prim `prim`."""
    out = np.zeros(len(prim), bool)
    p64 = arr["prims64"]
    ok = prim < len(p64)
    if ok.any():
        q = p64[prim[ok].astype(np.int64)]
        tri = (arr["prims"]["tf"][prim[ok].astype(np.int64)] & 15) == 0
        bc = bary(point[ok], q[:, 0:3], q[:, 3:6], q[:, 6:9])
        out[np.flatnonzero(ok)] = tri & (np.abs(bc).min(1) < tol)
    return out


class HipBuffers:
    """Device buffers through the HIP runtime the library itself links (the
    test allocates; the entry point under test does not)."""

    def __init__(self):
        xa.lib()  # loads libxtracer_amd.so and with it the runtime
        h = ctypes.CDLL("libamdhip64.so.7")
        h.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        h.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        h.hipFree.argtypes = [ctypes.c_void_p]
        self.h = h
        self.held = []

    def put(self, a):
        a = np.ascontiguousarray(a)
        p = ctypes.c_void_p()
        assert self.h.hipMalloc(ctypes.byref(p), max(a.nbytes, 4)) == 0
        assert self.h.hipMemcpy(p, a.ctypes.data, a.nbytes, 1) == 0  # host to device
        self.held.append(p)
        return p

    def get(self, p, like):
        out = np.empty_like(like)
        assert self.h.hipMemcpy(out.ctypes.data, p, out.nbytes, 2) == 0  # device to host
        return out

    def free(self):
        for p in self.held:
            self.h.hipFree(p)
        self.held = []


def same_stream_check(img, ref, mean, m2, spp, min_close, z=4.0, max_far=1e-3):
    """Device image against the oracle's on the same random streams: at least
    `min_close` of the pixels within 1e-3 relative, and at most `max_far` of
    the channels beyond z standard errors of the difference of two
    independent spp-sample estimates (a path that took another FP32 decision
    is another sample; with a few samples a pixel whose samples happened to
    agree can see such a sample far out). The per-sample variance comes from
    the oracle's own samples (mean, m2); where they were all equal, from an
    event of probability 3 / spp changing the value by its size (not seen in
    spp samples with 95% confidence). Returns the fraction of close pixels."""
    close = np.isclose(img[..., :3], ref[..., :3], rtol=1e-3, atol=1e-4).all(axis=2)
    var = m2 / max(1, spp - 1)
    var = np.where(var > 0, var, np.maximum(mean ** 2, 1e-6) * 3.0 / spp)
    bound = z * np.sqrt(2.0 * var / spp) + 1e-4
    dev = np.abs(img[..., :3].astype(np.float64) - ref[..., :3])
    far = dev > bound
    assert close.mean() >= min_close, f"only {close.mean():.4f} of pixels agree"
    assert far.mean() <= max_far, f"{far.sum()} channels beyond {z} standard errors, e.g. {np.argwhere(far)[:3].tolist()}"
    return close.mean()
