"""ctypes binding of libxtracer_amd.so (the C ABI declared in include/xtracer_amd.h).

The library is built in-tree by ``make -C xtracer_amd`` (or ``__graft_entry__.build()``).
Nothing here falls back to a CPU path: if the library is missing, loading raises.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libxtracer_amd.so")

XT_OK = 0
ERRORS = {1: "XT_E_INVALID", 2: "XT_E_IO", 3: "XT_E_SCENE", 4: "XT_E_DEVICE", 5: "XT_E_BUDGET",
          6: "XT_E_NOTFOUND", 7: "XT_E_STATE"}

ACCEL_KDTREE, ACCEL_OCTREE = 0, 1
STENCIL, DEPTH, AO, PATHTRACER, NORMAL, UV, EMISSION, WIREFRAME = range(8)
INTEGRATORS = {"stencil": STENCIL, "depth": DEPTH, "ao": AO, "pathtracer": PATHTRACER, "normal": NORMAL, "uv": UV,
               "emission": EMISSION, "wireframe": WIREFRAME}
AA_GRID, AA_RANDOM = 0, 1
ACCUM_DEFAULT, ACCUM_REFERENCE, ACCUM_FP64 = 0, 1, 2
HIT_NONE = 0xFFFFFFFF
HIT_PLANE_BASE = 0xFFFF0000
TUNING_AUTO = 0xFFFFFFFF  # xt_tuning.trace_blocks_per_cu: resident fill
SERVE_FAIR, SERVE_SRPT, SERVE_FCFS, SERVE_HYBRID = 0, 1, 2, 3
SERVE_POLICIES = {"fair": SERVE_FAIR, "srpt": SERVE_SRPT, "fcfs": SERVE_FCFS, "hybrid": SERVE_HYBRID}
CAMERA_THIN_LENS, CAMERA_ODS, CAMERA_ERP, CAMERA_CUBEMAP = 0, 1, 2, 3

# Exported symbols, in the order include/xtracer_amd.h declares them.
SYMBOLS = [
    "xt_scene_load", "xt_scene_build", "xt_scene_free", "xt_scene_error", "xt_scene_get_info",
    "xt_scene_get_object", "xt_scene_get_arrays", "xt_scene_get_object_params", "xt_scene_get_vnormals", "xt_scene_get_texcoords", "xt_scene_get_samplers",
    "xt_scene_get_kdtree", "xt_scene_build_kd_device", "xt_scene_get_kd_device", "xt_scene_mesh_octree", "xt_scene_camera", "xt_camera_make",
    "xt_image_write", "xt_image_read", "xt_render_raygraph", "xt_dev_create", "xt_dev_destroy", "xt_last_error", "xt_dev_set_tuning", "xt_dev_get_tuning",
    "xt_scene_upload", "xt_scene_release", "xt_trace",
    "xt_trace_device", "xt_trace_device_static", "xt_shade", "xt_render", "xt_render_device", "xt_submit", "xt_poll", "xt_serve_step", "xt_drain",
    "xt_serve_get_info", "xt_serve_set_policy", "xt_serve_configure", "xt_serve_get_config",
]


class XtError(RuntimeError):
    def __init__(self, code, msg=""):
        self.code = code
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")


class BuildParams(ctypes.Structure):
    _fields_ = [("kd_max_depth", ctypes.c_int), ("kd_leaf_prims", ctypes.c_int),
                ("kd_cost_intersect", ctypes.c_float), ("threads", ctypes.c_int), ("kd_on_device", ctypes.c_int)]


class KdBuildStats(ctypes.Structure):
    _fields_ = [("ms", ctypes.c_double)] + [(n, ctypes.c_uint32) for n in (
        "n_nodes", "n_refs", "depth", "leaves", "levels", "pad")]

    def as_dict(self):
        return {n: getattr(self, n) for n, _ in self._fields_ if n != "pad"}


class SceneInfo(ctypes.Structure):
    _fields_ = [(n, ctypes.c_uint32) for n in (
        "n_objects", "n_prims", "n_planes", "n_materials", "n_cameras", "n_kd_nodes", "n_kd_refs", "kd_depth",
        "n_oct_nodes", "n_oct_refs", "env_type", "any_vertex_normals")] + [
        ("build_seconds_octree", ctypes.c_double), ("build_seconds_kd", ctypes.c_double)]


class ObjectInfo(ctypes.Structure):
    _fields_ = [("name", ctypes.c_char * 64), ("handle", ctypes.c_uint64), ("geom_type", ctypes.c_uint32),
                ("material", ctypes.c_uint32), ("first_prim", ctypes.c_uint32), ("prim_count", ctypes.c_uint32),
                ("plane", ctypes.c_int32)]


class Camera(ctypes.Structure):
    _fields_ = [("position", ctypes.c_double * 3), ("basis", ctypes.c_double * 9), ("dir_z", ctypes.c_double),
                ("aperture", ctypes.c_double), ("flength", ctypes.c_double), ("thin_lens", ctypes.c_uint32),
                ("kind", ctypes.c_uint32), ("ipd", ctypes.c_double)]


class Tuning(ctypes.Structure):
    _fields_ = [(n, ctypes.c_uint32) for n in ("kd_stack", "trace_block", "wave_paths", "trace_blocks_per_cu",
                                                "trace_segments", "ray_order", "kd_layout", "kd_leaf",
                                                "frame_lanes")]

    def as_dict(self):
        return {n: getattr(self, n) for n, _ in self._fields_}


class ServeConfig(ctypes.Structure):
    _fields_ = [(n, ctypes.c_uint32) for n in ("regen", "graph", "rounds_per_step", "paths", "sample_buffer_mb",
                                                "lane_streams", "loop")]

    def as_dict(self):
        return {n: getattr(self, n) for n, _ in self._fields_}


class ServeInfo(ctypes.Structure):
    _fields_ = [("queries", ctypes.c_uint64), ("retests", ctypes.c_uint64), ("samples", ctypes.c_uint64),
                ("passes", ctypes.c_uint64), ("jobs_waiting", ctypes.c_uint32), ("jobs_active", ctypes.c_uint32),
                ("jobs_finished", ctypes.c_uint32), ("pad", ctypes.c_uint32), ("ms_device", ctypes.c_double),
                ("bytes_held", ctypes.c_uint64)]

    def as_dict(self):
        return {n: getattr(self, n) for n, _ in self._fields_ if n != "pad"}


class Job(ctypes.Structure):
    _fields_ = [(n, ctypes.c_uint32) for n in (
        "scene_id", "integrator", "width", "height", "samples", "aa", "aa_mode", "rdepth", "accel", "accum",
        "seed", "job_key", "samples_per_pass", "pad")] + [("camera", Camera)]


class JobStats(ctypes.Structure):
    _fields_ = [("queries", ctypes.c_uint64), ("retests", ctypes.c_uint64), ("samples", ctypes.c_uint64),
                ("ms_total", ctypes.c_double), ("ms_trace", ctypes.c_double), ("trace_launches", ctypes.c_uint32),
                ("passes", ctypes.c_uint32), ("latency_ms", ctypes.c_double), ("queue_ms", ctypes.c_double),
                ("trace_fail", ctypes.c_uint64), ("retest_fail", ctypes.c_uint64), ("ms_trace_busy", ctypes.c_double)]

    def as_dict(self):
        return {n: getattr(self, n) for n, _ in self._fields_}


class Rays(ctypes.Structure):
    _fields_ = [(n, ctypes.POINTER(ctypes.c_float)) for n in ("ox", "oy", "oz", "dx", "dy", "dz")]


class Hits(ctypes.Structure):
    _fields_ = [("t", ctypes.POINTER(ctypes.c_float)), ("prim", ctypes.POINTER(ctypes.c_uint32)),
                ("u", ctypes.POINTER(ctypes.c_float)), ("v", ctypes.POINTER(ctypes.c_float)),
                ("flags", ctypes.POINTER(ctypes.c_uint32))]


# numpy views of the flattened arrays (xtracer_amd/csrc/xt_layout.h)
PRIM_DTYPE = np.dtype([("p0", "<f4", 3), ("tf", "<u4"), ("p1", "<f4", 3), ("object", "<u4"), ("p2", "<f4", 3),
                       ("aux", "<f4")])
PLANE_DTYPE = np.dtype([("normal", "<f4", 3), ("object", "<u4"), ("point", "<f4", 3), ("pad", "<u4"),
                        ("normal64", "<f8", 3), ("point64", "<f8", 3)])
SHADE_QUERY_DTYPE = np.dtype([("point", "<f8", 3), ("normal", "<f8", 3), ("texcoord", "<f8", 3), ("camera", "<f8", 3),
                              ("emitter", "<f8", 3), ("intensity", "<f4", 3), ("object", "<u4")])
MATERIAL_DTYPE = np.dtype([("type", "<u4"), ("kd", "<f4", 3), ("ks", "<f4", 3), ("ke", "<f4", 3),
                           ("reflectance", "<f4"), ("exponent", "<f4")])

_lib = None


def lib():
    """Loads libxtracer_amd.so once; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: build it with `make -C xtracer_amd` "
                          "(or __graft_entry__.build()); there is no CPU fallback")
    L = ctypes.CDLL(LIB_PATH)
    P, U32, U64, I = ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint64, ctypes.c_int
    sig = {
        "xt_scene_load": (I, [ctypes.c_char_p, ctypes.POINTER(ctypes.c_char_p), I, ctypes.POINTER(P)]),
        "xt_scene_build": (I, [P, ctypes.POINTER(BuildParams)]),
        "xt_scene_free": (None, [P]),
        "xt_scene_error": (ctypes.c_char_p, [P]),
        "xt_scene_get_info": (I, [P, ctypes.POINTER(SceneInfo)]),
        "xt_scene_get_object": (I, [P, U32, ctypes.POINTER(ObjectInfo)]),
        "xt_scene_get_arrays": (I, [P, P, P, P, P, P, P]),
        "xt_scene_get_object_params": (I, [P, U32, P]),
        "xt_scene_get_vnormals": (I, [P, P]),
        "xt_scene_get_texcoords": (I, [P, P, P]),
        "xt_scene_get_samplers": (I, [P, P, P, P, P, P, P]),
        "xt_scene_get_kdtree": (I, [P, P, P, P]),
        "xt_scene_build_kd_device": (I, [P, U32, ctypes.POINTER(BuildParams), ctypes.POINTER(KdBuildStats)]),
        "xt_scene_get_kd_device": (I, [P, U32, P, P, P, ctypes.POINTER(KdBuildStats)]),
        "xt_scene_mesh_octree": (I, [P, U32, ctypes.POINTER(U64), ctypes.POINTER(U64), ctypes.POINTER(U64),
                                     ctypes.POINTER(U64)]),
        "xt_scene_camera": (I, [P, ctypes.c_char_p, ctypes.POINTER(Camera)]),
        "xt_camera_make": (I, [ctypes.POINTER(ctypes.c_double)] * 3 + [ctypes.c_float] * 3 +
                           [ctypes.POINTER(Camera)]),
        "xt_image_write": (I, [ctypes.c_char_p, P, U32, U32]),
        "xt_image_read": (I, [ctypes.c_char_p, P, P, P]),
        "xt_render_raygraph": (I, [P, ctypes.POINTER(Job), P, ctypes.c_char_p]),
        "xt_dev_create": (I, [ctypes.POINTER(P), I, U64]),
        "xt_dev_destroy": (None, [P]),
        "xt_last_error": (ctypes.c_char_p, [P]),
        "xt_dev_set_tuning": (I, [P, ctypes.POINTER(Tuning)]),
        "xt_dev_get_tuning": (I, [P, ctypes.POINTER(Tuning)]),
        "xt_scene_upload": (I, [P, P, ctypes.POINTER(U32)]),
        "xt_scene_release": (I, [P, U32]),
        "xt_trace": (I, [P, U32, I, ctypes.POINTER(Rays), ctypes.POINTER(Hits), U64]),
        "xt_trace_device": (I, [P, U32, I, P, P, P, P, U64, ctypes.POINTER(ctypes.c_float)]),
        "xt_trace_device_static": (I, [P, U32, I, P, P, P, P, U64]),
        "xt_shade": (I, [P, U32, P, U64, P, P]),
        "xt_render": (I, [P, ctypes.POINTER(Job), P, ctypes.POINTER(JobStats)]),
        "xt_render_device": (I, [P, ctypes.POINTER(Job), P, ctypes.POINTER(JobStats)]),
        "xt_submit": (I, [P, ctypes.POINTER(Job), ctypes.POINTER(U64)]),
        "xt_poll": (I, [P, U64, ctypes.POINTER(I), P, ctypes.POINTER(JobStats)]),
        "xt_serve_step": (I, [P, ctypes.POINTER(I)]),
        "xt_drain": (I, [P]),
        "xt_serve_get_info": (I, [P, ctypes.POINTER(ServeInfo)]),
        "xt_serve_set_policy": (I, [P, I]),
        "xt_serve_configure": (I, [P, ctypes.POINTER(ServeConfig)]),
        "xt_serve_get_config": (I, [P, ctypes.POINTER(ServeConfig)]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


class Scene:
    """Host scene: io::scn::load + flattening + octree/kd-tree builds (no GPU needed)."""

    def __init__(self, path, mods=(), build=True, **build_params):
        L = lib()
        self._h = ctypes.c_void_p()
        arr = (ctypes.c_char_p * max(1, len(mods)))(*[m.encode() for m in mods])
        rc = L.xt_scene_load(str(path).encode(), arr, len(mods), ctypes.byref(self._h))
        if rc != XT_OK:
            msg = L.xt_scene_error(self._h).decode() if self._h else ""
            if self._h:
                L.xt_scene_free(self._h)
                self._h = None
            raise XtError(rc, msg)
        if build:
            self.build(**build_params)

    def build(self, kd_max_depth=0, kd_leaf_prims=0, kd_cost_intersect=0.0, threads=0, kd_on_device=False):
        """kd_on_device: leave the kd-tree to the device, which builds it when
        the scene is uploaded."""
        p = BuildParams(kd_max_depth, kd_leaf_prims, kd_cost_intersect, threads, int(bool(kd_on_device)))
        rc = lib().xt_scene_build(self._h, ctypes.byref(p))
        if rc != XT_OK:
            raise XtError(rc, lib().xt_scene_error(self._h).decode())

    @property
    def handle(self):
        return self._h

    def info(self):
        i = SceneInfo()
        rc = lib().xt_scene_get_info(self._h, ctypes.byref(i))
        if rc:
            raise XtError(rc)
        return i

    def objects(self):
        out = []
        for k in range(self.info().n_objects):
            o = ObjectInfo()
            rc = lib().xt_scene_get_object(self._h, k, ctypes.byref(o))
            if rc:
                raise XtError(rc)
            out.append({"name": o.name.decode(), "handle": o.handle, "geom_type": o.geom_type,
                        "material": o.material, "first_prim": o.first_prim, "prim_count": o.prim_count,
                        "plane": o.plane})
        return out

    def arrays(self):
        i = self.info()
        prims = np.zeros(i.n_prims, PRIM_DTYPE)
        prims64 = np.zeros((i.n_prims, 9), np.float64)
        local = np.zeros(i.n_prims, np.uint32)
        planes = np.zeros(i.n_planes, PLANE_DTYPE)
        mats = np.zeros(i.n_materials, MATERIAL_DTYPE)
        env = np.zeros(7, np.float32)
        rc = lib().xt_scene_get_arrays(self._h, _ptr(prims), _ptr(prims64), _ptr(local), _ptr(planes), _ptr(mats),
                                       _ptr(env))
        if rc:
            raise XtError(rc)
        return {"prims": prims, "prims64": prims64, "prim_local": local, "planes": planes, "materials": mats,
                "env": env}

    def object_params(self, index):
        out = np.zeros(9, np.float64)
        rc = lib().xt_scene_get_object_params(self._h, index, _ptr(out))
        if rc:
            raise XtError(rc)
        return out

    def vnormals(self):
        out = np.zeros((self.info().n_prims, 9), np.float32)
        rc = lib().xt_scene_get_vnormals(self._h, _ptr(out))
        if rc == 6:
            return None
        if rc:
            raise XtError(rc)
        return out

    def texcoords(self):
        """(prim_tc: (n_prims, 6) u, v per triangle corner; uv_scale: (n_objects, 2))."""
        i = self.info()
        tc = np.zeros((i.n_prims, 6), np.float64)
        sc = np.zeros((len(self.objects()), 2), np.float64)
        rc = lib().xt_scene_get_texcoords(self._h, _ptr(tc), _ptr(sc))
        if rc:
            raise XtError(rc)
        return tc, sc

    def samplers(self):
        """Textures and table samplers: dict(tex_info (n, 4) uint32, texels (m, 4) float32,
        samplers (k, 14) uint32 words, mat_smp (n_materials, 3), env_smp)."""
        counts = np.zeros(3, np.uint32)
        self._check_rc(lib().xt_scene_get_samplers(self._h, _ptr(counts), None, None, None, None, None))
        nt, nf, ns = (int(c) for c in counts)
        tex = np.zeros((nt, 4), np.uint32)
        texels = np.zeros((nf // 4, 4), np.float32)
        smp = np.zeros((ns, 14), np.uint32)
        ms = np.zeros((self.info().n_materials, 3), np.uint32)
        env = np.zeros(1, np.uint32)
        self._check_rc(lib().xt_scene_get_samplers(self._h, _ptr(counts), _ptr(tex), _ptr(texels), _ptr(smp), _ptr(ms),
                                                   _ptr(env)))
        return {"tex_info": tex, "texels": texels, "samplers": smp, "mat_smp": ms, "env_smp": int(env[0])}

    def _check_rc(self, rc):
        if rc:
            raise XtError(rc)

    def kdtree(self):
        i = self.info()
        nodes = np.zeros((i.n_kd_nodes, 2), np.uint32)
        refs = np.zeros(i.n_kd_refs, np.uint32)
        bounds = np.zeros(6, np.float32)
        rc = lib().xt_scene_get_kdtree(self._h, _ptr(nodes), _ptr(refs), _ptr(bounds))
        if rc:
            raise XtError(rc)
        return nodes, refs, bounds

    def mesh_octrees(self):
        res = []
        k = 0
        while True:
            g, n, lv, nt = (ctypes.c_uint64() for _ in range(4))
            rc = lib().xt_scene_mesh_octree(self._h, k, ctypes.byref(g), ctypes.byref(n), ctypes.byref(lv),
                                            ctypes.byref(nt))
            if rc:
                return res
            res.append({"geometry": g.value, "nodes": n.value, "max_level": lv.value, "triangles": nt.value})
            k += 1

    def camera(self, name=None):
        c = Camera()
        rc = lib().xt_scene_camera(self._h, (name or "").encode(), ctypes.byref(c))
        if rc:
            raise XtError(rc, f"camera {name!r}")
        return c

    def close(self):
        if self._h:
            lib().xt_scene_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def make_camera(position, target, up, fov, aperture=0.0, flength=0.0):
    c = Camera()
    d3 = ctypes.c_double * 3
    rc = lib().xt_camera_make(d3(*position), d3(*target), d3(*up), fov, aperture, flength, ctypes.byref(c))
    if rc:
        raise XtError(rc)
    return c


def write_image(path, rgba):
    """Writes an (H, W, 4) float framebuffer to `path`; the format follows the
    extension (.png .bmp .tga .hdr .exr .pfm .ppm). Host-only: needs the
    library, not a GPU."""
    import numpy as np
    a = np.ascontiguousarray(rgba, dtype=np.float32)
    if a.ndim != 3 or a.shape[2] != 4:
        raise ValueError(f"expected an (H, W, 4) array, got {a.shape}")
    rc = lib().xt_image_write(os.fsencode(path), a.ctypes.data, a.shape[1], a.shape[0])
    if rc:
        raise XtError(rc, f"xt_image_write({path})")


def read_image(path):
    """(H, W, 4) float32 pixels of a texture image, as scenes load them
    (xt_image_read). Host-only."""
    import numpy as np
    w = ctypes.c_uint32()
    h = ctypes.c_uint32()
    rc = lib().xt_image_read(os.fsencode(path), None, ctypes.byref(w), ctypes.byref(h))
    if rc:
        raise XtError(rc, f"xt_image_read({path})")
    a = np.zeros((h.value, w.value, 4), np.float32)
    rc = lib().xt_image_read(os.fsencode(path), a.ctypes.data, ctypes.byref(w), ctypes.byref(h))
    if rc:
        raise XtError(rc, f"xt_image_read({path})")
    return a


class Device:
    """One HIP device with resident scenes (xt_dev)."""

    def __init__(self, device=0, hbm_budget=0):
        L = lib()
        self._d = ctypes.c_void_p()
        rc = L.xt_dev_create(ctypes.byref(self._d), device, hbm_budget)
        if rc != XT_OK:
            raise XtError(rc, "xt_dev_create")

    def _check(self, rc):
        if rc != XT_OK:
            raise XtError(rc, lib().xt_last_error(self._d).decode())

    @property
    def handle(self):
        return self._d

    def tuning(self):
        t = Tuning()
        self._check(lib().xt_dev_get_tuning(self._d, ctypes.byref(t)))
        return t.as_dict()

    def set_tuning(self, **knobs):
        """Sets launch-shape knobs (kd_stack, trace_block, wave_paths,
        trace_blocks_per_cu, trace_segments, ray_order, kd_layout, kd_leaf,
        frame_lanes); unnamed knobs keep their value."""
        t = Tuning(**{k: int(v) for k, v in knobs.items()})
        self._check(lib().xt_dev_set_tuning(self._d, ctypes.byref(t)))
        return self.tuning()

    def upload(self, scene):
        sid = ctypes.c_uint32()
        self._check(lib().xt_scene_upload(self._d, scene.handle, ctypes.byref(sid)))
        return sid.value

    def release(self, scene_id):
        self._check(lib().xt_scene_release(self._d, scene_id))

    def build_kd(self, scene_id, kd_max_depth=0, kd_leaf_prims=0, kd_cost_intersect=0.0):
        """Rebuilds the resident scene's kd-tree on the device (the host
        build's tree, node for node); returns the build's stats."""
        p = BuildParams(kd_max_depth, kd_leaf_prims, kd_cost_intersect, 0, 0)
        st = KdBuildStats()
        self._check(lib().xt_scene_build_kd_device(self._d, scene_id, ctypes.byref(p), ctypes.byref(st)))
        return st.as_dict()

    def kdtree(self, scene_id):
        """The resident kd-tree: nodes (n, 2) uint32, refs, bounds (6,), stats."""
        st = KdBuildStats()
        self._check(lib().xt_scene_get_kd_device(self._d, scene_id, None, None, None, ctypes.byref(st)))
        nodes = np.zeros((max(st.n_nodes, 1), 2), np.uint32)
        refs = np.zeros(max(st.n_refs, 1), np.uint32)
        bounds = np.zeros(6, np.float32)
        self._check(lib().xt_scene_get_kd_device(self._d, scene_id, _ptr(nodes), _ptr(refs), _ptr(bounds),
                                                 ctypes.byref(st)))
        return nodes[:st.n_nodes], refs[:st.n_refs], bounds, st.as_dict()

    def trace(self, scene_id, origins, directions, accel=ACCEL_KDTREE):
        o = np.ascontiguousarray(origins, np.float32)
        d = np.ascontiguousarray(directions, np.float32)
        n = o.shape[0]
        cols = [np.ascontiguousarray(o[:, k]) for k in range(3)] + [np.ascontiguousarray(d[:, k]) for k in range(3)]
        fp = ctypes.POINTER(ctypes.c_float)
        rays = Rays(*[c.ctypes.data_as(fp) for c in cols])
        t = np.zeros(n, np.float32)
        prim = np.zeros(n, np.uint32)
        u = np.zeros(n, np.float32)
        v = np.zeros(n, np.float32)
        flags = np.zeros(n, np.uint32)
        hits = Hits(t.ctypes.data_as(fp), prim.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)),
                    u.ctypes.data_as(fp), v.ctypes.data_as(fp), flags.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)))
        self._check(lib().xt_trace(self._d, scene_id, accel, ctypes.byref(rays), ctypes.byref(hits), n))
        return t, prim, u, v, flags

    def render(self, job):
        img = np.zeros((job.height, job.width, 4), np.float32)
        st = JobStats()
        self._check(lib().xt_render(self._d, ctypes.byref(job), _ptr(img), ctypes.byref(st)))
        return img, st

    def render_raygraph(self, job, ply_path):
        """render() without stats, and the render's ray graph written to
        `ply_path` (xt_render_raygraph); returns the image."""
        img = np.zeros((job.height, job.width, 4), np.float32)
        self._check(lib().xt_render_raygraph(self._d, ctypes.byref(job), _ptr(img), os.fsencode(ply_path)))
        return img
    def render_device(self, job, out_ptr):
        st = JobStats()
        self._check(lib().xt_render_device(self._d, ctypes.byref(job), ctypes.c_void_p(out_ptr), ctypes.byref(st)))
        return st

    def trace_device(self, scene_id, ray0_ptr, ray1_ptr, hit_ptr, n, accel=ACCEL_KDTREE, flags_ptr=None):
        ms = ctypes.c_float()
        self._check(lib().xt_trace_device(self._d, scene_id, accel, ctypes.c_void_p(ray0_ptr),
                                          ctypes.c_void_p(ray1_ptr), ctypes.c_void_p(hit_ptr),
                                          ctypes.c_void_p(flags_ptr) if flags_ptr else None, n, ctypes.byref(ms)))
        return ms.value

    def shade(self, scene_id, q, emitter, intensity):
        """IMaterial::shade on the device (see xt_shade). q: (n, 16) float64
        rows of hit point, normal, texcoord, camera position, emitter
        position, object index; emitter, intensity: (n, 3) float32. Returns
        (intensity after shade, results)."""
        q = np.asarray(q, np.float64)
        rec = np.zeros(len(q), SHADE_QUERY_DTYPE)
        rec["point"], rec["normal"], rec["texcoord"] = q[:, 0:3], q[:, 3:6], q[:, 6:9]
        rec["camera"], rec["emitter"] = q[:, 9:12], q[:, 12:15]
        rec["intensity"] = np.asarray(emitter, np.float32)
        rec["object"] = q[:, 15].astype(np.uint32)
        io = np.array(intensity, np.float32, copy=True, order="C")
        ret = np.zeros(len(q), np.int32)
        self._check(lib().xt_shade(self._d, scene_id, _ptr(rec), len(q), _ptr(io), _ptr(ret)))
        return io, ret

    def submit(self, job):
        t = ctypes.c_uint64()
        self._check(lib().xt_submit(self._d, ctypes.byref(job), ctypes.byref(t)))
        return t.value

    def poll(self, ticket, width, height):
        """(done, image or None, JobStats); a finished ticket is forgotten."""
        done = ctypes.c_int()
        st = JobStats()
        self._check(lib().xt_poll(self._d, ticket, ctypes.byref(done), None, ctypes.byref(st)))
        if not done.value:
            return False, None, st
        img = np.empty((height, width, 4), np.float32)
        self._check(lib().xt_poll(self._d, ticket, ctypes.byref(done), _ptr(img), ctypes.byref(st)))
        return True, img, st

    def set_serve_policy(self, policy):
        """SERVE_FAIR / SERVE_SRPT / SERVE_FCFS / SERVE_HYBRID (or their names)."""
        if isinstance(policy, str):
            policy = SERVE_POLICIES[policy]
        self._check(lib().xt_serve_set_policy(self._d, int(policy)))

    def serve_configure(self, regen=None, graph=None, rounds_per_step=0, paths=0, sample_buffer_mb=0,
                        lane_streams=None, loop=None):
        """regen / graph / lane_streams / loop (device-fed serving): True or
        False (None keeps the value); the counts: 0 keeps the value. Returns
        the configuration in force."""
        def flag(v):
            return 0 if v is None else (2 if v else 1)
        c = ServeConfig(flag(regen), flag(graph), int(rounds_per_step), int(paths), int(sample_buffer_mb),
                        flag(lane_streams), flag(loop))
        self._check(lib().xt_serve_configure(self._d, ctypes.byref(c)))
        return self.serve_config()

    def serve_config(self):
        c = ServeConfig()
        self._check(lib().xt_serve_get_config(self._d, ctypes.byref(c)))
        return c.as_dict()

    def serve_info(self):
        info = ServeInfo()
        self._check(lib().xt_serve_get_info(self._d, ctypes.byref(info)))
        return info.as_dict()

    def serve_step(self):
        n = ctypes.c_int()
        self._check(lib().xt_serve_step(self._d, ctypes.byref(n)))
        return n.value

    def drain(self):
        self._check(lib().xt_drain(self._d))

    def close(self):
        if self._d:
            lib().xt_dev_destroy(self._d)
            self._d = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def make_job(scene_id, integrator, width, height, camera, samples=1, aa=1, aa_mode=AA_GRID, rdepth=100,
             accel=ACCEL_KDTREE, accum=ACCUM_DEFAULT, seed=1, job_key=0, samples_per_pass=0):
    """An xt_job; defaults follow params_t (src/xtcore/context.cc:3-27)."""
    if isinstance(integrator, str):
        integrator = INTEGRATORS[integrator]
    return Job(scene_id, integrator, width, height, samples, aa, aa_mode, rdepth, accel, accum, seed, job_key,
               samples_per_pass, 0, camera)
