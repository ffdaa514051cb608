"""Tiny .scn writer for hand-built test scenes."""
import numpy as np

from xtracer_amd.scenes import col3, f32, vec3


def write(path, geometry, objects, materials, camera=None, env=None):
    """geometry: {name: dict(type=..., **props)}; objects: {name: (geometry, material)};
    materials: {name: dict(type=..., kd=, ks=, ke=, reflectance=, exponent=)};
    camera: dict(position, target, fov, aperture, flength), or dict(type=ods|erp|cubemap,
    position, orientation, ipd); env: None | ("color", c) | ("gradient", a, b) | a sampler dict.
    A material channel (kd / ks / ke) is a colour or a sampler dict: dict(type=texture, source,
    filtering, flip_x, flip_y), dict(type=erp, source), dict(type=cubemap, posx=.., negz=..),
    dict(type=gradient, a, b)."""
    out = ["title = test"]
    if isinstance(env, dict):
        out.append("environment = {\n\ttype = %s\n\tconfig = {\n%s\n\t}\n}"
                   % (env["type"], "\n".join(_sampler_props(env, "\t\t"))))
    elif env is not None:
        if env[0] == "color":
            out.append("environment = {\n\ttype = color\n\tconfig = {\n\t\tvalue = %s\n\t}\n}" % col3(env[1]))
        else:
            out.append("environment = {\n\ttype = gradient\n\tconfig = {\n\t\ta = %s\n\t\tb = %s\n\t}\n}"
                       % (col3(env[1]), col3(env[2])))
    cam = camera or dict(position=(0, 0, -5), target=(0, 0, 0), fov=45.0)
    kind = cam.get("type", "thin-lens")
    if kind == "thin-lens":
        out.append("camera = {\n\tdefault = {\n\t\ttype = thin-lens\n\t\tfov = %s\n\t\taperture = %s\n\t\tflength = %s\n"
                   "\t\tposition = %s\n\t\ttarget = %s\n\t\tup = vec3(0,1,0)\n\t}\n}"
                   % (f32(cam.get("fov", 45.0)), f32(cam.get("aperture", 0.0)), f32(cam.get("flength", 0.0)),
                      vec3(cam["position"]), vec3(cam["target"])))
    else:  # ods / erp / cubemap (parseutil.cc:222-258)
        c = [f"\t\ttype = {kind}", f"\t\tposition = {vec3(cam['position'])}"]
        if "orientation" in cam:
            c.append(f"\t\torientation = {vec3(cam['orientation'])}")
        if "ipd" in cam:
            c.append(f"\t\tipd = {f32(cam['ipd'])}")
        out.append("camera = {\n\tdefault = {\n" + "\n".join(c) + "\n\t}\n}")
    g = ["geometry = {"]
    for name, p in geometry.items():
        g.append(f"\t{name} = {{\n\t\ttype = {p['type']}")
        if p["type"] == "plane":
            g.append(f"\t\tnormal = {vec3(p['normal'])}\n\t\tdistance = {p['distance']!r}")
        elif p["type"] in ("sphere", "point"):
            g.append(f"\t\tposition = {vec3(p['position'])}")
            if p["type"] == "sphere":
                g.append(f"\t\tradius = {p['radius']!r}")
        elif p["type"] == "triangle" and "tc" in p:  # corners as groups with u, v (parseutil.cc:385-397)
            g.append("\t\tvecdata = {")
            for k, (v, t) in enumerate(zip(p["v"], p["tc"])):
                g.append("\t\t\tv%d = {\n\t\t\t\tx = %r\n\t\t\t\ty = %r\n\t\t\t\tz = %r\n\t\t\t\tu = %r\n\t\t\t\tv = %r\n\t\t\t}"
                         % (k, float(v[0]), float(v[1]), float(v[2]), float(t[0]), float(t[1])))
            g.append("\t\t}")
        elif p["type"] == "triangle":
            g.append("\t\tvecdata = {\n\t\t\tv0 = %s\n\t\t\tv1 = %s\n\t\t\tv2 = %s\n\t\t}" % tuple(vec3(v) for v in p["v"]))
        elif p["type"] == "mesh":
            g.append(f"\t\tsource = {p['source']}")
        for k in ("u_scale", "v_scale"):
            if k in p:
                g.append(f"\t\t{k} = {p[k]!r}")
        g.append("\t}")
    g.append("}")
    out.append("\n".join(g))
    o = ["object = {"]
    for name, (geo, mat) in objects.items():
        o.append(f"\t{name} = {{\n\t\tgeometry = {geo}\n\t\tmaterial = {mat}\n\t}}")
    o.append("}")
    out.append("\n".join(o))
    m = ["material = {"]
    for name, p in materials.items():
        m.append(f"\t{name} = {{\n\t\ttype = {p['type']}\n\t\tproperties = {{\n\t\t\tsamplers = {{")
        for key, sname in (("kd", "diffuse"), ("ks", "specular"), ("ke", "emissive")):
            if key in p and isinstance(p[key], dict):
                m.append(f"\t\t\t\t{sname} = {{\n\t\t\t\t\ttype = {p[key]['type']}\n"
                         + "\n".join(_sampler_props(p[key], "\t\t\t\t\t")) + "\n\t\t\t\t}")
            elif key in p:
                m.append(f"\t\t\t\t{sname} = {{\n\t\t\t\t\ttype = color\n\t\t\t\t\tvalue = {col3(p[key])}\n\t\t\t\t}}")
        m.append("\t\t\t}")
        sc = [(k, p[k]) for k in ("reflectance", "exponent", "ior") if k in p]
        if sc:
            m.append("\t\t\tscalars = {")
            m += [f"\t\t\t\t{k} = {f32(v)}" for k, v in sc]
            m.append("\t\t\t}")
        m.append("\t\t}\n\t}")
    m.append("}")
    out.append("\n".join(m))
    with open(path, "w") as f:
        f.write("\n".join(out) + "\n")
    return str(path)


def _sampler_props(smp, ind):
    """Property lines of a sampler dict (parseutil.cc:147-203, 472-500)."""
    lines = []
    for k, v in smp.items():
        if k == "type":
            continue
        if k in ("a", "b", "value"):
            v = col3(v)
        elif isinstance(v, bool):
            v = "true" if v else "false"
        lines.append(f"{ind}{k} = {v}")
    return lines


def write_obj(path, tris):
    tris = np.asarray(tris, np.float32).reshape(-1, 3, 3)
    with open(path, "w") as f:
        for t in tris:
            for v in t:
                f.write("v %.9g %.9g %.9g\n" % tuple(np.float64(v)))
        for k in range(len(tris)):
            f.write("f %d %d %d\n" % (3 * k + 1, 3 * k + 2, 3 * k + 3))
    return str(path)


def write_obj_uv(path, tris, uvs):
    """OBJ with texture coordinates: uvs[k] is None (corners without vt) or
    three (u, v) pairs for triangle k."""
    tris = np.asarray(tris, np.float32).reshape(-1, 3, 3)
    with open(path, "w") as f:
        for t in tris:
            for v in t:
                f.write("v %.9g %.9g %.9g\n" % tuple(np.float64(v)))
        nt = 0
        for uv in uvs:
            if uv is not None:
                for u, v in uv:
                    f.write("vt %.9g %.9g\n" % (u, v))
        for k, uv in enumerate(uvs):
            if uv is None:
                f.write("f %d %d %d\n" % (3 * k + 1, 3 * k + 2, 3 * k + 3))
            else:
                f.write("f %d/%d %d/%d %d/%d\n" % (3 * k + 1, nt + 1, 3 * k + 2, nt + 2, 3 * k + 3, nt + 3))
                nt += 3
    return str(path)
