"""Seeded synthetic scenes in the reference's .scn format (plus OBJ meshes).

Every coordinate is written with %.9g from a float32 value, so the
reference-style sscanf("%f") parse reproduces it exactly. Generated scenes are
cached under a directory keyed by their parameters.

  spheres   64 Lambert spheres (centres U[0,10]^3, radii U[0.2,1]) on a plane,
            an inert point light, pinhole camera, gradient environment
  icosphere icosphere (20 * 4^subdiv triangles) radially displaced by 3-octave
            value noise (amplitude 5% of the radius) on a plane, an emissive
            box mesh and an emissive sphere, Lambert/Phong, thin lens
  strands   random-walk strands of thin triangles inside the unit sphere, an
            emissive mesh light, Blinn-Phong
  terrain   value-noise heightfield
"""
import hashlib
import json
import os

import numpy as np

DEFAULT_CACHE = os.environ.get("XTRACER_SCENE_CACHE", os.path.join("/tmp", "xtracer_amd_scenes"))

GRADIENT_ENV = """environment = {
    type = gradient
    config = {
        a = col3(1,1,1)
        b = col3(0.5,0.7,1)
    }
}
"""


def f32(x):
    return "%.9g" % float(np.float32(x))


def vec3(v):
    return "vec3(%s,%s,%s)" % tuple(f32(c) for c in v)


def col3(v):
    return "col3(%s,%s,%s)" % tuple(f32(c) for c in v)


def _cache_dir(kind, params, cache):
    key = hashlib.sha1(json.dumps([kind, params], sort_keys=True).encode()).hexdigest()[:12]
    d = os.path.join(cache or DEFAULT_CACHE, f"{kind}_{key}")
    os.makedirs(d, exist_ok=True)
    return d


def _done(path):
    return os.path.exists(path) and os.path.exists(path + ".done")


def _mark(path):
    with open(path + ".done", "w") as f:
        f.write("ok\n")


_CHUNK = 1 << 20


def _fmt_rows(pattern, rows):
    """Formats rows of 3 numbers with one %-pattern per row, a million rows
    per string operation. Everything runs in this process: a scene may be
    generated after the GPU runtime has been initialised (by torch, by the
    native library or by a profiler's preload), and forking then is unsafe."""
    out = []
    for k in range(0, len(rows), _CHUNK):
        part = rows[k:k + _CHUNK]
        out.append((pattern * len(part)) % tuple(part.ravel().tolist()))
    return "".join(out)


def write_obj(path, verts, faces, normals=None):
    """Writes vertices (float32-exact) and 1-based triangle faces."""
    verts = np.asarray(verts, np.float32)
    faces = np.asarray(faces, np.int64)
    tmp = path + ".tmp"
    with open(tmp, "w") as f:
        f.write("# xtracer_amd synthetic mesh\n")
        f.write(_fmt_rows("v %.9g %.9g %.9g\n", verts.astype(np.float64)))
        if normals is not None:
            normals = np.asarray(normals, np.float32).astype(np.float64)
            f.write("".join("vn %.9g %.9g %.9g\n" % tuple(n) for n in normals))
            f.write("".join("f %d//%d %d//%d %d//%d\n" % (a, a, b, b, c, c) for a, b, c in (faces + 1)))
        else:
            f.write(_fmt_rows("f %d %d %d\n", faces + 1))
    os.replace(tmp, path)


def _material(name, kind, kd=None, ks=None, ke=None, reflectance=None, exponent=None):
    lines = [f"\t{name} = {{", f"\t\ttype = {kind}", "\t\tproperties = {", "\t\t\tsamplers = {"]
    for sname, val in (("diffuse", kd), ("specular", ks), ("emissive", ke)):
        if val is not None:
            lines += [f"\t\t\t\t{sname} = {{", "\t\t\t\t\ttype = color", f"\t\t\t\t\tvalue = {col3(val)}",
                      "\t\t\t\t}"]
    lines.append("\t\t\t}")
    if reflectance is not None or exponent is not None:
        lines.append("\t\t\tscalars = {")
        if reflectance is not None:
            lines.append(f"\t\t\t\treflectance = {f32(reflectance)}")
        if exponent is not None:
            lines.append(f"\t\t\t\texponent = {f32(exponent)}")
        lines.append("\t\t\t}")
    lines += ["\t\t}", "\t}"]
    return "\n".join(lines)


def _camera(position, target, fov=45.0, aperture=0.0, flength=0.0):
    return ("camera = {\n\tdefault = {\n\t\ttype = thin-lens\n"
            f"\t\tfov = {f32(fov)}\n\t\taperture = {f32(aperture)}\n\t\tflength = {f32(flength)}\n"
            f"\t\tposition = {vec3(position)}\n\t\ttarget = {vec3(target)}\n\t\tup = vec3(0,1,0)\n\t}}\n}}\n")


def camera_dict(position, target, fov=45.0, aperture=0.0, flength=0.0):
    """The camera in the form the oracle takes (float32-rounded like the loader)."""
    r = lambda v: [float(np.float32(c)) for c in v]  # noqa: E731
    return {"position": r(position), "target": r(target), "up": [0.0, 1.0, 0.0], "fov": float(np.float32(fov)),
            "aperture": float(np.float32(aperture)), "flength": float(np.float32(flength))}


# ------------------------------------------------------------------ spheres

SPHERES_CAMERA = dict(position=(5.0, 6.0, -14.0), target=(5.0, 4.0, 5.0), fov=45.0)


def spheres(seed=0, count=64, cache=None):
    """Plane + `count` Lambert spheres + a point light; returns the .scn path."""
    d = _cache_dir("spheres", [seed, count], cache)
    path = os.path.join(d, "scene.scn")
    if _done(path):
        return path
    rng = np.random.default_rng(seed)
    c = rng.uniform(0.0, 10.0, (count, 3)).astype(np.float32)
    r = rng.uniform(0.2, 1.0, count).astype(np.float32)
    kd = rng.uniform(0.2, 0.8, (count, 3)).astype(np.float32)
    geo = ["geometry = {", "\tfloor = {\n\t\ttype = plane\n\t\tnormal = vec3(0,1,0)\n\t\tdistance = 0\n\t}",
           "\tlamp = {\n\t\ttype = point\n\t\tposition = vec3(5,15,5)\n\t}"]
    obj = ["object = {", "\tfloor = {\n\t\tgeometry = floor\n\t\tmaterial = grey\n\t}",
           "\tlamp = {\n\t\tgeometry = lamp\n\t\tmaterial = light\n\t}"]
    mats = ["material = {", _material("grey", "lambert", kd=(0.5, 0.5, 0.5)),
            _material("light", "emissive", ke=(1, 1, 1))]
    for i in range(count):
        geo.append(f"\ts{i:03d} = {{\n\t\ttype = sphere\n\t\tposition = {vec3(c[i])}\n\t\tradius = {f32(r[i])}\n\t}}")
        obj.append(f"\tball{i:03d} = {{\n\t\tgeometry = s{i:03d}\n\t\tmaterial = m{i:03d}\n\t}}")
        mats.append(_material(f"m{i:03d}", "lambert", kd=kd[i]))
    text = "title = spheres\n" + GRADIENT_ENV + _camera(**SPHERES_CAMERA) + "\n".join(geo) + "\n}\n" + \
        "\n".join(obj) + "\n}\n" + "\n".join(mats) + "\n}\n"
    with open(path, "w") as f:
        f.write(text)
    _mark(path)
    return path


# ------------------------------------------------------------------ icosphere

def icosphere(subdiv):
    """Unit icosphere: (vertices float64, faces int64), outward winding."""
    t = (1.0 + 5 ** 0.5) / 2.0
    v = np.array([[-1, t, 0], [1, t, 0], [-1, -t, 0], [1, -t, 0], [0, -1, t], [0, 1, t], [0, -1, -t], [0, 1, -t],
                  [t, 0, -1], [t, 0, 1], [-t, 0, -1], [-t, 0, 1]], np.float64)
    f = np.array([[0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11], [1, 5, 9], [5, 11, 4], [11, 10, 2],
                  [10, 7, 6], [7, 1, 8], [3, 9, 4], [3, 4, 2], [3, 2, 6], [3, 6, 8], [3, 8, 9], [4, 9, 5],
                  [2, 4, 11], [6, 2, 10], [8, 6, 7], [9, 8, 1]], np.int64)
    v /= np.linalg.norm(v, axis=1)[:, None]
    for _ in range(subdiv):
        e = np.concatenate([f[:, [0, 1]], f[:, [1, 2]], f[:, [2, 0]]])
        lo, hi = np.minimum(e[:, 0], e[:, 1]), np.maximum(e[:, 0], e[:, 1])
        key = lo * len(v) + hi
        uk, inv = np.unique(key, return_inverse=True)
        a, b = uk // len(v), uk % len(v)
        mid = v[a] + v[b]
        mid /= np.linalg.norm(mid, axis=1)[:, None]
        m = len(v) + inv.reshape(3, -1).T  # midpoints of edges 01, 12, 20
        v = np.concatenate([v, mid])
        f = np.concatenate([np.stack([f[:, 0], m[:, 0], m[:, 2]], 1), np.stack([f[:, 1], m[:, 1], m[:, 0]], 1),
                            np.stack([f[:, 2], m[:, 2], m[:, 1]], 1), m])
    return v, f


def value_noise(p, seed, octaves=3, base_freq=2.0):
    """Seeded 3-D value noise in [-1, 1]: lattice values from an integer hash,
    smoothstep trilinear interpolation, frequencies f, 2f, 4f with amplitudes
    1, 1/2, 1/4."""
    def lattice(ix, iy, iz):
        h = (ix * 73856093) ^ (iy * 19349663) ^ (iz * 83492791) ^ (seed * 2654435761)
        h = (h ^ (h >> 13)) * 1274126177
        h = h ^ (h >> 16)
        return ((h & 0xFFFFFF).astype(np.float64) / float(0xFFFFFF)) * 2.0 - 1.0

    total = np.zeros(len(p))
    amp, freq, norm = 1.0, base_freq, 0.0
    for _ in range(octaves):
        q = p * freq
        i0 = np.floor(q).astype(np.int64)
        fr = q - i0
        s = fr * fr * (3 - 2 * fr)
        acc = np.zeros(len(p))
        for dx in (0, 1):
            for dy in (0, 1):
                for dz in (0, 1):
                    w = (s[:, 0] if dx else 1 - s[:, 0]) * (s[:, 1] if dy else 1 - s[:, 1]) * \
                        (s[:, 2] if dz else 1 - s[:, 2])
                    acc += w * lattice(i0[:, 0] + dx, i0[:, 1] + dy, i0[:, 2] + dz)
        total += amp * acc
        norm += amp
        amp *= 0.5
        freq *= 2.0
    return total / norm


def box_mesh(center, size):
    cx, cy, cz = center
    sx, sy, sz = (s / 2 for s in size)
    v = np.array([[cx + x * sx, cy + y * sy, cz + z * sz] for x in (-1, 1) for y in (-1, 1) for z in (-1, 1)])
    f = np.array([[0, 1, 3], [0, 3, 2], [4, 6, 7], [4, 7, 5], [0, 4, 5], [0, 5, 1], [2, 3, 7], [2, 7, 6],
                  [0, 2, 6], [0, 6, 4], [1, 5, 7], [1, 7, 3]])
    return v, f


ICOSPHERE_CAMERA = dict(position=(0.0, 1.6, 5.0), target=(0.0, 1.0, 0.0), fov=45.0, aperture=0.05, flength=2.5)


def icosphere_scene(seed=0, subdiv=8, cache=None):
    """Displaced icosphere on a plane with a box light and a sphere light."""
    d = _cache_dir("icosphere", [seed, subdiv], cache)
    path = os.path.join(d, "scene.scn")
    if _done(path):
        return path
    rng = np.random.default_rng(seed)
    v, f = icosphere(subdiv)
    r = 1.0 + 0.05 * value_noise(v, seed)
    v = v * r[:, None] + np.array([0.0, 1.06, 0.0])
    write_obj(os.path.join(d, "icosphere.obj"), v, f)
    bv, bf = box_mesh((0.0, 3.2, 0.0), (1.2, 0.1, 1.2))
    write_obj(os.path.join(d, "boxlight.obj"), bv, bf)
    kd_s = rng.uniform(0.2, 0.8, 3)
    kd_p = rng.uniform(0.2, 0.8, 3)
    text = ("title = icosphere\n" + GRADIENT_ENV + _camera(**ICOSPHERE_CAMERA) +
            "geometry = {\n"
            "\tfloor = {\n\t\ttype = plane\n\t\tnormal = vec3(0,1,0)\n\t\tdistance = 0\n\t}\n"
            "\tico = {\n\t\ttype = mesh\n\t\tsource = icosphere.obj\n\t}\n"
            "\tbox = {\n\t\ttype = mesh\n\t\tsource = boxlight.obj\n\t}\n"
            "\tbulb = {\n\t\ttype = sphere\n\t\tposition = vec3(2,2.2,1.5)\n\t\tradius = 0.35\n\t}\n}\n"
            "object = {\n"
            "\tfloor = {\n\t\tgeometry = floor\n\t\tmaterial = floor_phong\n\t}\n"
            "\tico = {\n\t\tgeometry = ico\n\t\tmaterial = ico_lambert\n\t}\n"
            "\tbox = {\n\t\tgeometry = box\n\t\tmaterial = box_light\n\t}\n"
            "\tbulb = {\n\t\tgeometry = bulb\n\t\tmaterial = bulb_light\n\t}\n}\n"
            "material = {\n" +
            _material("floor_phong", "phong", kd=kd_p, ks=(0.6, 0.6, 0.6), reflectance=0.25, exponent=30) + "\n" +
            _material("ico_lambert", "lambert", kd=kd_s) + "\n" +
            _material("box_light", "emissive", ke=(6, 6, 6)) + "\n" +
            _material("bulb_light", "emissive", ke=(8, 6, 4)) + "\n}\n")
    with open(path, "w") as fh:
        fh.write(text)
    _mark(path)
    return path


# ------------------------------------------------------------------ strands

STRANDS_CAMERA = dict(position=(0.0, 0.3, 2.6), target=(0.0, 0.0, 0.0), fov=50.0)


def strands_scene(seed=0, strands=50000, points=60, step=0.05, width=0.002, cache=None):
    """Random-walk strands (one thin triangle per segment) in the unit sphere."""
    d = _cache_dir("strands", [seed, strands, points, step, width], cache)
    path = os.path.join(d, "scene.scn")
    if _done(path):
        return path
    rng = np.random.default_rng(seed)
    p = np.zeros((strands, points, 3))
    start = rng.normal(size=(strands, 3))
    start /= np.linalg.norm(start, axis=1)[:, None]
    p[:, 0] = start * rng.uniform(0, 1, (strands, 1)) ** (1 / 3) * 0.95
    for k in range(1, points):
        s = rng.normal(size=(strands, 3))
        s /= np.linalg.norm(s, axis=1)[:, None]
        q = p[:, k - 1] + step * s
        n = np.linalg.norm(q, axis=1)
        q[n > 0.98] *= (0.98 / n[n > 0.98])[:, None]
        p[:, k] = q
    seg = p[:, 1:] - p[:, :-1]
    side = np.cross(seg, np.array([0.0, 1.0, 0.3]))
    side /= np.maximum(np.linalg.norm(side, axis=2), 1e-12)[:, :, None]
    a = p[:, :-1] - 0.5 * width * side
    b = p[:, :-1] + 0.5 * width * side
    c = p[:, 1:]
    verts = np.stack([a, b, c], axis=2).reshape(-1, 3)
    faces = np.arange(len(verts)).reshape(-1, 3)
    write_obj(os.path.join(d, "strands.obj"), verts, faces)
    lv, lf = box_mesh((0.0, 1.6, 0.0), (1.0, 0.05, 1.0))
    write_obj(os.path.join(d, "light.obj"), lv, lf)
    text = ("title = strands\n" + GRADIENT_ENV + _camera(**STRANDS_CAMERA) +
            "geometry = {\n"
            "\thair = {\n\t\ttype = mesh\n\t\tsource = strands.obj\n\t}\n"
            "\tpanel = {\n\t\ttype = mesh\n\t\tsource = light.obj\n\t}\n}\n"
            "object = {\n"
            "\thair = {\n\t\tgeometry = hair\n\t\tmaterial = hair_bp\n\t}\n"
            "\tpanel = {\n\t\tgeometry = panel\n\t\tmaterial = panel_light\n\t}\n}\n"
            "material = {\n" +
            _material("hair_bp", "blinn_phong", kd=(0.55, 0.35, 0.2), ks=(0.5, 0.5, 0.5), reflectance=0.3,
                      exponent=40) + "\n" +
            _material("panel_light", "emissive", ke=(5, 5, 5)) + "\n}\n")
    with open(path, "w") as fh:
        fh.write(text)
    _mark(path)
    return path


# ------------------------------------------------------------------ terrain

TERRAIN_CAMERA = dict(position=(0.0, 2.5, 5.5), target=(0.0, 0.0, 0.0), fov=50.0)


def terrain_scene(seed=0, n=317, cache=None):
    """n x n vertex value-noise heightfield over [-4, 4]^2 (2 (n-1)^2 triangles)."""
    d = _cache_dir("terrain", [seed, n], cache)
    path = os.path.join(d, "scene.scn")
    if _done(path):
        return path
    xs = np.linspace(-4.0, 4.0, n)
    X, Z = np.meshgrid(xs, xs)
    P = np.stack([X.ravel(), np.zeros(X.size), Z.ravel()], 1)
    Y = 0.6 * value_noise(P * 0.4 + 0.13, seed, octaves=3)
    verts = np.stack([X.ravel(), Y, Z.ravel()], 1)
    i = np.arange(n - 1)
    I, J = np.meshgrid(i, i)
    a = (J * n + I).ravel()
    faces = np.concatenate([np.stack([a, a + n, a + 1], 1), np.stack([a + 1, a + n, a + n + 1], 1)])
    write_obj(os.path.join(d, "terrain.obj"), verts, faces)
    text = ("title = terrain\n" + GRADIENT_ENV + _camera(**TERRAIN_CAMERA) +
            "geometry = {\n\tland = {\n\t\ttype = mesh\n\t\tsource = terrain.obj\n\t}\n"
            "\tsun = {\n\t\ttype = sphere\n\t\tposition = vec3(2,4,-2)\n\t\tradius = 0.6\n\t}\n}\n"
            "object = {\n\tland = {\n\t\tgeometry = land\n\t\tmaterial = land_lambert\n\t}\n"
            "\tsun = {\n\t\tgeometry = sun\n\t\tmaterial = sun_light\n\t}\n}\n"
            "material = {\n" + _material("land_lambert", "lambert", kd=(0.45, 0.5, 0.35)) + "\n" +
            _material("sun_light", "emissive", ke=(6, 6, 5)) + "\n}\n")
    with open(path, "w") as fh:
        fh.write(text)
    _mark(path)
    return path


SCENES = {"spheres": (spheres, SPHERES_CAMERA), "icosphere": (icosphere_scene, ICOSPHERE_CAMERA),
          "strands": (strands_scene, STRANDS_CAMERA), "terrain": (terrain_scene, TERRAIN_CAMERA)}
