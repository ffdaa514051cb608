"""TEST INFRASTRUCTURE ONLY: Python access to the oracle (oracle/liboracle.so).

Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg use this
module, and only as the checker.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
REF_LIB = os.path.join(HERE, "_ref", "libxtref.so")

_lib = None
_ref = None


class SceneDesc(ctypes.Structure):
    P = ctypes.c_void_p
    _fields_ = [("n_objects", ctypes.c_uint32), ("obj_type", P), ("obj_material", P), ("obj_param", P),
                ("mesh_first", P), ("mesh_count", P), ("mesh_v", P), ("mesh_n", P), ("obj_tc", P), ("mesh_tc", P),
                ("obj_uv_scale", P), ("n_materials", ctypes.c_uint32),
                ("mat_type", P), ("mat_param", P), ("env_type", ctypes.c_uint32), ("env", P),
                ("n_textures", ctypes.c_uint32), ("tex_info", P), ("texels", P), ("n_samplers", ctypes.c_uint32),
                ("samplers", P), ("mat_smp", P), ("env_smp", ctypes.c_uint32)]


class OrJob(ctypes.Structure):
    _fields_ = [(n, ctypes.c_uint32) for n in ("integrator", "width", "height", "samples", "aa", "aa_mode",
                                                "rdepth", "seed", "job_key", "tile_size", "tile_stride")] + [
        ("threads", ctypes.c_int), ("cam", ctypes.c_double * 9), ("fov", ctypes.c_float),
        ("aperture", ctypes.c_float), ("flength", ctypes.c_float), ("cam_kind", ctypes.c_uint32),
        ("orientation", ctypes.c_double * 3), ("ipd", ctypes.c_float)]


CAMERA_KINDS = {"thin-lens": 0, "ods": 1, "erp": 2, "cubemap": 3}


def _job(integrator, width, height, cam, samples=1, aa=1, aa_mode=0, rdepth=100, seed=1, job_key=0, threads=0,
         tile_stride=1):
    """cam: dict(position, target, up, fov, aperture, flength) for a thin lens, or
    dict(type=ods|erp|cubemap, position, orientation, ipd)."""
    kind = CAMERA_KINDS[cam.get("type", "thin-lens")]
    return OrJob(integrator, width, height, samples, aa, aa_mode, rdepth, seed, job_key, 32, tile_stride, threads,
                 (ctypes.c_double * 9)(*cam["position"], *cam.get("target", (0, 0, 1)), *cam.get("up", (0, 1, 0))),
                 cam.get("fov", 0.0), cam.get("aperture", 0.0), cam.get("flength", 0.0), kind,
                 (ctypes.c_double * 3)(*cam.get("orientation", (0, 0, 0))), cam.get("ipd", 0.0))


def camera_rays(cam, width, height, xy):
    """The oracle's primary rays of camera dict `cam` at sample coordinates xy (n x 2 float32)."""
    xy = np.ascontiguousarray(xy, np.float32)
    o = np.zeros((len(xy), 3))
    d = np.zeros((len(xy), 3))
    lib().or_camera_rays(ctypes.byref(_job(0, width, height, cam)), len(xy), _p(xy), _p(o), _p(d))
    return o, d


def thin_lens_rays(cam, width, height, xy, jitter):
    """The oracle's thin-lens rays at sample coordinates xy (n x 2 float32)
    with the lens jitters given (n x 2 doubles, the values prng_c returned)."""
    xy = np.ascontiguousarray(xy, np.float32)
    jitter = np.ascontiguousarray(jitter, np.float64)
    o = np.zeros((len(xy), 3))
    d = np.zeros((len(xy), 3))
    lib().or_thin_lens_rays(ctypes.byref(_job(0, width, height, cam)), len(xy), _p(xy), _p(jitter), _p(o), _p(d))
    return o, d


def lens_jitter(cam, keys):
    """The lens jitters (n x 2) the oracle draws for sample keys `keys`."""
    keys = np.ascontiguousarray(keys, np.uint32)
    out = np.zeros((len(keys), 2))
    lib().or_lens_jitter(ctypes.byref(_job(0, 1, 1, cam)), len(keys), _p(keys), _p(out))
    return out


def aa_random(width, level, pixels, seed=1, job_key=0, draws=None):
    """RANDOM anti-aliasing of `pixels` (n x 2 uint32 x, y) at `level`: the
    float sample coordinates (n, level, 2) and the uniform draws (n, level, 2)
    they come from; with `draws` given those are used instead of the
    oracle's own."""
    pixels = np.ascontiguousarray(pixels, np.uint32)
    n = len(pixels)
    coords = np.zeros((n, level, 2), np.float32)
    out = np.zeros((n, level, 2))
    cam = {"position": (0, 0, 0)}
    jb = _job(0, width, 1, cam, aa=level, aa_mode=1, seed=seed, job_key=job_key)
    dr = None if draws is None else np.ascontiguousarray(draws, np.float64)
    lib().or_aa_random(ctypes.byref(jb), n, _p(pixels), _p(dr), _p(out), _p(coords))
    return coords, out


def path_key(seed, job_key, pixel, sample):
    return int(lib().or_path_key(seed, job_key, pixel, sample))


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            raise ImportError(f"{LIB} missing: build it with `make -C oracle`")
        L = ctypes.CDLL(LIB)
        P = ctypes.c_void_p
        L.or_build.restype = P
        L.or_build.argtypes = [ctypes.POINTER(SceneDesc)]
        L.or_free.argtypes = [P]
        L.or_trace.argtypes = [P, ctypes.c_uint64, P, P, P, P, P, P, ctypes.c_int]
        L.or_render.restype = ctypes.c_uint64
        L.or_render.argtypes = [P, ctypes.POINTER(OrJob), P, P, P]
        L.or_camera_rays.argtypes = [ctypes.POINTER(OrJob), ctypes.c_uint64, P, P, P]
        L.or_thin_lens_rays.argtypes = [ctypes.POINTER(OrJob), ctypes.c_uint64, P, P, P, P]
        L.or_lens_jitter.argtypes = [ctypes.POINTER(OrJob), ctypes.c_uint64, P, P]
        L.or_aa_random.argtypes = [ctypes.POINTER(OrJob), ctypes.c_uint64, P, P, P, P]
        L.or_path_key.restype = ctypes.c_uint32
        L.or_path_key.argtypes = [ctypes.c_uint32] * 4
        L.or_mesh_octree.argtypes = [P, ctypes.c_uint32, ctypes.POINTER(ctypes.c_uint64), P]
        L.or_shade.argtypes = [P, ctypes.c_uint64, P, P, P, P]
        L.or_sample.restype = ctypes.c_uint64
        L.or_sample.argtypes = [ctypes.c_int, P, P, ctypes.c_double, ctypes.c_uint32, ctypes.c_uint64, P]
        _lib = L
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


class OracleScene:
    """Oracle scene built from a product Scene's export: objects in scan order,
    mesh triangles put back into OBJ face order with prim_local."""

    def __init__(self, scene):
        objs = scene.objects()
        arr = scene.arrays()
        vn = scene.vnormals()
        n = len(objs)
        self.obj_type = np.array([o["geom_type"] for o in objs], np.uint32)
        self.obj_material = np.array([o["material"] for o in objs], np.uint32)
        self.obj_param = np.zeros((n, 9), np.float64)
        self.mesh_first = np.zeros(n, np.uint32)
        self.mesh_count = np.zeros(n, np.uint32)
        tris, norms, tcs = [], [], []
        prim_tc, uv_scale = scene.texcoords()
        self.obj_tc = np.zeros((n, 6), np.float64)
        self.obj_uv_scale = np.ascontiguousarray(uv_scale, np.float64)
        base = 0
        # soup index -> (object, OBJ-order index) map, for comparing hits
        self.soup_to_local = arr["prim_local"].astype(np.int64)
        for k, o in enumerate(objs):
            self.obj_param[k] = scene.object_params(k)
            if o["geom_type"] == 2:
                self.obj_tc[k] = prim_tc[o["first_prim"]]
            if o["geom_type"] == 3:
                lo, cnt = o["first_prim"], o["prim_count"]
                order = np.argsort(arr["prim_local"][lo:lo + cnt], kind="stable")
                tris.append(arr["prims64"][lo:lo + cnt][order])
                tcs.append(prim_tc[lo:lo + cnt][order])
                if vn is not None:
                    norms.append(vn[lo:lo + cnt][order])
                self.mesh_first[k] = base
                self.mesh_count[k] = cnt
                base += cnt
        self.mesh_v = np.ascontiguousarray(np.concatenate(tris) if tris else np.zeros((1, 9)), np.float64)
        self.mesh_n = (np.ascontiguousarray(np.concatenate(norms), np.float32) if (vn is not None and norms) else None)
        self.mesh_tc = np.ascontiguousarray(np.concatenate(tcs) if tcs else np.zeros((1, 6)), np.float64)
        mats = arr["materials"]
        self.mat_type = np.ascontiguousarray(mats["type"], np.uint32)
        mp = np.zeros((max(1, len(mats)), 11), np.float32)
        if len(mats):
            mp[:, 0:3] = mats["kd"]
            mp[:, 3:6] = mats["ks"]
            mp[:, 6:9] = mats["ke"]
            mp[:, 9] = mats["reflectance"]
            mp[:, 10] = mats["exponent"]
        self.mat_param = mp
        self.env = np.ascontiguousarray(arr["env"][1:7], np.float32)
        smp = scene.samplers()
        self.tex_info = np.ascontiguousarray(smp["tex_info"])
        self.texels = np.ascontiguousarray(smp["texels"])
        self.samplers = np.ascontiguousarray(smp["samplers"])
        self.mat_smp = np.ascontiguousarray(smp["mat_smp"]) if len(mats) else None
        d = SceneDesc(n, _p(self.obj_type), _p(self.obj_material), _p(self.obj_param), _p(self.mesh_first),
                      _p(self.mesh_count), _p(self.mesh_v), _p(self.mesh_n), _p(self.obj_tc), _p(self.mesh_tc),
                      _p(self.obj_uv_scale), len(mats), _p(self.mat_type),
                      _p(self.mat_param), int(arr["env"][0]), _p(self.env), len(self.tex_info),
                      _p(self.tex_info), _p(self.texels), len(self.samplers), _p(self.samplers), _p(self.mat_smp),
                      smp["env_smp"])
        self._h = lib().or_build(ctypes.byref(d))
        self.objects = objs

    def trace(self, origins, directions, threads=0):
        n = len(origins)
        rays = np.ascontiguousarray(np.concatenate([origins, directions], axis=1), np.float64)
        t = np.zeros(n, np.float64)
        obj = np.zeros(n, np.int32)
        prim = np.zeros(n, np.int64)
        point = np.zeros((n, 3), np.float64)
        normal = np.zeros((n, 3), np.float64)
        lib().or_trace(self._h, n, _p(rays), _p(t), _p(obj), _p(prim), _p(point), _p(normal), threads)
        return t, obj, prim, point, normal

    def render(self, integrator, width, height, cam, samples=1, aa=1, aa_mode=0, rdepth=100, seed=1, job_key=0,
               threads=0, tile_stride=1, stats=False):
        """cam: a camera dict as _job takes it."""
        j = _job(integrator, width, height, cam, samples, aa, aa_mode, rdepth, seed, job_key, threads, tile_stride)
        img = np.zeros((height, width, 4), np.float32)
        mean = np.zeros((height, width, 3), np.float64) if stats else None
        m2 = np.zeros((height, width, 3), np.float64) if stats else None
        nq = lib().or_render(self._h, ctypes.byref(j), _p(img), _p(mean), _p(m2))
        if stats:
            return img, nq, mean, m2
        return img, nq

    def shade(self, q, emitter, intensity):
        """IMaterial::shade per query. q: (n, 16) float64 rows of hit point,
        normal, texcoord, camera position, emitter position, object index;
        emitter, intensity: (n, 3) float32. Returns (intensity after shade,
        shade()'s results: 1 / 0, -1 for an object without a material)."""
        q = np.ascontiguousarray(q, np.float64)
        io = np.array(intensity, np.float32, copy=True, order="C")
        ret = np.zeros(len(q), np.int32)
        lib().or_shade(self._h, len(q), _p(q), _p(np.ascontiguousarray(emitter, np.float32)), _p(io), _p(ret))
        return io, ret

    def mesh_preorder(self, obj):
        n = ctypes.c_uint64()
        cnt = int(self.mesh_count[obj])
        out = np.zeros(cnt, np.uint64)
        if lib().or_mesh_octree(self._h, obj, ctypes.byref(n), _p(out)):
            raise ValueError("not a mesh object")
        return n.value, out

    def __del__(self):
        try:
            if self._h:
                lib().or_free(self._h)
        except Exception:
            pass


def sample(kind, normal, count, seed=0, direction=(0.0, 0.0, 1.0), exponent=1.0):
    """Directions from the oracle's hemisphere / diffuse / lobe samplers."""
    kinds = {"hemisphere": 0, "diffuse": 1, "lobe": 2}
    n = np.ascontiguousarray(normal, np.float64)
    d = np.ascontiguousarray(direction, np.float64)
    out = np.zeros((count, 3), np.float64)
    used = lib().or_sample(kinds[kind], _p(n), _p(d), exponent, seed, count, _p(out))
    return out, used


def ref_available():
    return os.path.exists(REF_LIB)


def ref():
    """The real reference (oracle/_ref/libxtref.so) built from /root/reference."""
    global _ref
    if _ref is None:
        if not ref_available():
            raise ImportError(f"{REF_LIB} missing: build it with `make -C oracle/ref`")
        L = ctypes.CDLL(REF_LIB)
        P = ctypes.c_void_p
        L.xr_load.restype = P
        L.xr_load.argtypes = [ctypes.c_char_p, ctypes.POINTER(ctypes.c_int)]
        L.xr_free.argtypes = [P]
        L.xr_object_count.restype = ctypes.c_uint32
        L.xr_object_count.argtypes = [P]
        L.xr_object.argtypes = [P, ctypes.c_uint32, ctypes.c_char_p, ctypes.c_int, ctypes.POINTER(ctypes.c_int), P]
        L.xr_trace.argtypes = [P, ctypes.c_uint64, P, P, P, P, P]
        L.xr_octree_trace.argtypes = [ctypes.c_uint64, P, ctypes.c_uint64, P, P, P, P]
        L.xr_render.argtypes = [P, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_uint, P]
        L.xr_save_image.argtypes = [ctypes.c_int, ctypes.c_char_p, P, ctypes.c_int, ctypes.c_int]
        L.xr_camera_rays.argtypes = [P, ctypes.c_uint64, P, ctypes.c_float, ctypes.c_float, P, P]
        L.xr_thin_lens_rays.argtypes = [P, ctypes.c_uint64, P, ctypes.c_float, ctypes.c_float, ctypes.c_uint, P, P, P]
        L.xr_aa_random.argtypes = [ctypes.c_uint64, P, ctypes.c_uint, ctypes.c_uint, P, P, P]
        L.xr_load_image.argtypes = [ctypes.c_char_p, P, P, P]
        L.xr_write_raygraph.argtypes = [P, ctypes.c_char_p]
        L.xr_shade.argtypes = [P, ctypes.c_uint64, P, P, P, P]
        _ref = L
    return _ref


REF_IMAGE_FORMATS = {"png": 0, "bmp": 1, "tga": 2, "hdr": 3, "exr": 4}


def ref_save_image(fmt, path, rgba):
    """Writes an (H, W, 4) float framebuffer through the reference's own
    nimg::io::save::<fmt> (lib/nimg/img.cc:72-318), via a Pixmap."""
    a = np.ascontiguousarray(rgba, np.float32)
    rc = ref().xr_save_image(REF_IMAGE_FORMATS[fmt], str(path).encode(), _p(a), a.shape[1], a.shape[0])
    if rc:
        raise RuntimeError(f"reference {fmt} writer returned {rc}")


def ref_load_image(path):
    """(H, W, 4) float32 pixels of `path` as the reference's own loader reads
    them (nimg::io::load::image, lib/nimg/img.cc:17-66), or None when it
    refuses the file."""
    w = np.zeros(1, np.int32)
    h = np.zeros(1, np.int32)
    if ref().xr_load_image(str(path).encode(), None, _p(w), _p(h)):
        return None
    a = np.zeros((int(h[0]), int(w[0]), 4), np.float32)
    if ref().xr_load_image(str(path).encode(), _p(a), _p(w), _p(h)):
        return None
    return a


class RefScene:
    """A scene loaded by the real reference loader (io::scn::load)."""

    def __init__(self, path):
        st = ctypes.c_int()
        self._h = ref().xr_load(str(path).encode(), ctypes.byref(st))
        if not self._h:
            raise RuntimeError(f"reference loader returned {st.value} for {path}")

    def objects(self):
        out = []
        for i in range(ref().xr_object_count(self._h)):
            name = ctypes.create_string_buffer(128)
            kind = ctypes.c_int()
            par = np.zeros(9, np.float64)
            ref().xr_object(self._h, i, name, 128, ctypes.byref(kind), _p(par))
            out.append({"name": name.value.decode(), "kind": kind.value, "param": par})
        return out

    def trace(self, origins, directions):
        n = len(origins)
        rays = np.ascontiguousarray(np.concatenate([origins, directions], axis=1), np.float64)
        t = np.zeros(n, np.float64)
        obj = np.zeros(n, np.int32)
        point = np.zeros((n, 3), np.float64)
        normal = np.zeros((n, 3), np.float64)
        ref().xr_trace(self._h, n, _p(rays), _p(t), _p(obj), _p(point), _p(normal))
        return t, obj, point, normal

    def camera_rays(self, width, height, xy):
        """Primary rays of the scene's first camera at sample coordinates xy (n x 2 float32)."""
        xy = np.ascontiguousarray(xy, np.float32)
        o = np.zeros((len(xy), 3))
        d = np.zeros((len(xy), 3))
        if ref().xr_camera_rays(self._h, len(xy), _p(xy), width, height, _p(o), _p(d)):
            raise RuntimeError("scene has no camera")
        return o, d

    def thin_lens_rays(self, width, height, xy, seed=1):
        """Thin-lens primary rays of the scene's first camera at xy (n x 2
        float32), ray i under srand(seed + i), with the two lens jitters each
        ray drew (n x 2): (origins, directions, jitters)."""
        xy = np.ascontiguousarray(xy, np.float32)
        n = len(xy)
        o, d, j = np.zeros((n, 3)), np.zeros((n, 3)), np.zeros((n, 2))
        if ref().xr_thin_lens_rays(self._h, n, _p(xy), width, height, seed, _p(o), _p(d), _p(j)):
            raise RuntimeError("scene has no thin-lens camera")
        return o, d, j

    def render(self, integrator, width, height, samples=1, aa=1, rdepth=100, threads=1, seed=1, aa_mode=0):
        img = np.zeros((height, width, 4), np.float32)
        rc = ref().xr_render(self._h, integrator, width, height, samples, aa, aa_mode, rdepth, threads, seed,
                             _p(img))
        if rc:
            raise RuntimeError("xr_render failed")
        return img


    def shade(self, q, emitter, intensity):
        """The reference materials' own shade() (OracleScene.shade's layout)."""
        q = np.ascontiguousarray(q, np.float64)
        io = np.array(intensity, np.float32, copy=True, order="C")
        ret = np.zeros(len(q), np.int32)
        ref().xr_shade(self._h, len(q), _p(q), _p(np.ascontiguousarray(emitter, np.float32)), _p(io), _p(ret))
        return io, ret

    def write_raygraph(self, path):
        """PLY of the last render's ray graph, written by the reference."""
        if ref().xr_write_raygraph(self._h, str(path).encode()):
            raise RuntimeError("reference raygraph::write failed")
    def __del__(self):
        try:
            if self._h:
                ref().xr_free(self._h)
        except Exception:
            pass


def ref_aa_random(pixels, level, seed=1):
    """antialiasing::gen_samples_random of the real reference for each pixel
    (n x 2, pixel i under srand(seed + i)): the sample coordinates (n, level,
    2, doubles), weights (n, level) and the prng_c(0, 1) draws (n, level, 2)
    they came from."""
    pixels = np.ascontiguousarray(pixels, np.float32)
    n = len(pixels)
    coords, weights, draws = np.zeros((n, level, 2)), np.zeros((n, level)), np.zeros((n, level, 2))
    if ref().xr_aa_random(n, _p(pixels), level, seed, _p(coords), _p(weights), _p(draws)):
        raise RuntimeError("xr_aa_random failed")
    return coords, weights, draws


def ref_octree_trace(tris, origins, directions):
    """Octree<Triangle>::intersection of the real reference over `tris`
    (n x 9, OBJ face order); returns (triangle index or -1, t)."""
    n = len(origins)
    rays = np.ascontiguousarray(np.concatenate([origins, directions], axis=1), np.float64)
    tris = np.ascontiguousarray(tris, np.float64)
    idx = np.zeros(n, np.int64)
    t = np.zeros(n, np.float64)
    ref().xr_octree_trace(len(tris), _p(tris), n, _p(rays), _p(idx), _p(t), None)
    return idx, t
