"""Image parity of device renders against the oracle.

stencil / depth: bit-exact (float RGBA) with the reference's per-sample
read-modify-write. ao / pathtracer: the device and the oracle draw from the
same counter-based streams, so most pixels agree to float rounding; the
statistical check compares each device pixel with an independent oracle
estimate within the Monte Carlo 99% confidence bound."""
import numpy as np
import pytest

import xtracer_amd as xa
from oracle.oracle import OracleScene
from xtracer_amd import scenes
from tests.helpers import same_stream_check, scene_path


def _setup(device, kind, **kw):
    s = xa.Scene(scene_path(kind, **kw))
    sid = device.upload(s)
    cam = scenes.camera_dict(**scenes.SCENES[kind][1])
    return s, sid, cam


@pytest.mark.gpu
@pytest.mark.parametrize("integrator", [xa.STENCIL, xa.DEPTH, xa.NORMAL, xa.EMISSION, xa.WIREFRAME])
@pytest.mark.parametrize("kind,kw,samples", [("spheres", {}, 16), ("icosphere", {"subdiv": 4}, 4), ("glow", {}, 4)])
@pytest.mark.parametrize("accel", [xa.ACCEL_KDTREE, xa.ACCEL_OCTREE])
def test_bit_exact(device, kind, kw, samples, integrator, accel, tmp_path):
    """The deterministic integrators (stencil, depth, normal, emission,
    wireframe) reproduce the oracle's image bit for bit."""
    if kind == "glow":
        from tests.helpers import glow_scene
        s = xa.Scene(glow_scene(str(tmp_path / "glow.scn")))
        sid = device.upload(s)
        cam = scenes.camera_dict(position=(0.3, 1.4, -4.5), target=(0.0, 0.6, 0.5), fov=50.0)
    else:
        s, sid, cam = _setup(device, kind, **kw)
    W, H = 96, 64
    job = xa.make_job(sid, integrator, W, H, s.camera(), samples=samples, seed=5, accel=accel)
    img, st = device.render(job)
    ref, nq = OracleScene(s).render(integrator, W, H, cam, samples=samples, seed=5)
    assert st.queries == nq
    diff = np.flatnonzero((img.view(np.uint32) != ref.view(np.uint32)).any(axis=2))
    assert len(diff) == 0, f"{len(diff)} pixels differ, first {diff[:5]}: {img.reshape(-1, 4)[diff[:3]]} vs {ref.reshape(-1, 4)[diff[:3]]}"


@pytest.mark.gpu
@pytest.mark.parametrize("integrator", [xa.STENCIL, xa.DEPTH])
@pytest.mark.parametrize("kind,kw", [("spheres", {}), ("icosphere", {"subdiv": 4})])
@pytest.mark.parametrize("aa", [1, 3])
def test_aa_random_bit_exact(device, kind, kw, integrator, aa):
    """RANDOM anti-aliasing (src/xtcore/aa.cc:34-48): the device places each
    pixel's sub-samples with the oracle's draws, so stencil and depth stay bit
    for bit the oracle's; the icosphere's camera is a thin lens."""
    s, sid, cam = _setup(device, kind, **kw)
    W, H, spp = 80, 56, 2
    job = xa.make_job(sid, integrator, W, H, s.camera(), samples=spp, aa=aa, aa_mode=xa.AA_RANDOM, seed=9)
    img, st = device.render(job)
    ref, nq = OracleScene(s).render(integrator, W, H, cam, samples=spp, aa=aa, aa_mode=xa.AA_RANDOM, seed=9)
    assert st.queries == nq == W * H * spp * aa
    diff = np.flatnonzero((img.view(np.uint32) != ref.view(np.uint32)).any(axis=2))
    assert len(diff) == 0, f"{len(diff)} pixels differ"
    # the sub-sample positions matter: the GRID image differs
    grid, _ = OracleScene(s).render(integrator, W, H, cam, samples=spp, aa=aa, seed=9)
    assert (aa == 1 and integrator == xa.STENCIL) or not np.array_equal(grid, ref)


@pytest.mark.gpu
@pytest.mark.parametrize("integrator", [xa.AO, xa.PATHTRACER])
def test_aa_random_same_stream(device, integrator):
    """RANDOM anti-aliasing with the stochastic integrators: same paths as the
    oracle except where an FP32 decision differs."""
    s, sid, cam = _setup(device, "spheres")
    W, H, spp = 48, 32, 4
    job = xa.make_job(sid, integrator, W, H, s.camera(), samples=spp, aa=2, aa_mode=xa.AA_RANDOM, rdepth=5,
                      seed=13, accum=xa.ACCUM_REFERENCE)
    img, st = device.render(job)
    ref, nq, mean, m2 = OracleScene(s).render(integrator, W, H, cam, samples=spp, aa=2, aa_mode=xa.AA_RANDOM, rdepth=5,
                                              seed=13, stats=True)
    same_stream_check(img, ref, mean, m2, spp * 2, 0.995)
    assert abs(st.queries - nq) <= 0.005 * nq


@pytest.mark.gpu
@pytest.mark.parametrize("integrator,kind,kw", [(xa.AO, "spheres", {}), (xa.PATHTRACER, "icosphere", {"subdiv": 4}),
                                                (xa.PATHTRACER, "spheres", {})])
def test_same_stream(device, integrator, kind, kw):
    """Same seed: device and oracle trace the same paths except where an FP32
    decision differs from FP64."""
    s, sid, cam = _setup(device, kind, **kw)
    W, H, spp = 64, 48, 8
    job = xa.make_job(sid, integrator, W, H, s.camera(), samples=spp, rdepth=6, seed=11, accum=xa.ACCUM_REFERENCE)
    img, st = device.render(job)
    ref, nq, mean, m2 = OracleScene(s).render(integrator, W, H, cam, samples=spp, rdepth=6, seed=11, stats=True)
    same_stream_check(img, ref, mean, m2, spp, 0.995)
    assert abs(st.queries - nq) <= 0.01 * nq


@pytest.mark.gpu
@pytest.mark.parametrize("integrator,kind,kw", [(xa.AO, "spheres", {}), (xa.PATHTRACER, "icosphere", {"subdiv": 4})])
def test_monte_carlo_bound(device, integrator, kind, kw):
    """Different seeds: per-pixel difference within the 99% MC confidence bound."""
    s, sid, cam = _setup(device, kind, **kw)
    W, H, spp = 32, 24, 64
    job = xa.make_job(sid, integrator, W, H, s.camera(), samples=spp, rdepth=6, seed=101)
    img, _ = device.render(job)
    _, _, mean, m2 = OracleScene(s).render(integrator, W, H, cam, samples=4 * spp, rdepth=6, seed=202, stats=True)
    var = m2 / max(1, 4 * spp - 1)
    se = np.sqrt(var / spp + var / (4 * spp))
    z = 2.576
    inside = np.abs(img[..., :3] - mean) <= z * se + 1e-6
    assert inside.mean() >= 0.97, f"{inside.mean():.4f} of channels inside the 99% bound"


@pytest.mark.gpu
@pytest.mark.parametrize("kind", ["pinhole", "erp", "cubemap"])
@pytest.mark.parametrize("accel", [xa.ACCEL_KDTREE, xa.ACCEL_OCTREE])
def test_uv(device, kind, accel, tmp_path):
    """The uv integrator: texture coordinates from triangle corners, mesh vt
    indices (and Triangle() defaults), scaled and wrapped plane projections
    and the spheres' asin mapping (NaN where asin's argument passes 1, as in
    the reference). Triangles and planes use only the basic
    operations, so their pixels must be bit-exact; a sphere's asin may round
    differently in the device's math library, so pixels must agree to 1e-6
    relative and all but 0.5% bit for bit."""
    from tests.helpers import camera_dict_of, uv_scene
    s = xa.Scene(uv_scene(str(tmp_path / "uv.scn"), kind))
    sid = device.upload(s)
    W, H = 96, 72
    job = xa.make_job(sid, xa.UV, W, H, s.camera(), samples=4, seed=5, accel=accel)
    img, st = device.render(job)
    device.release(sid)
    ref, nq = OracleScene(s).render(xa.UV, W, H, camera_dict_of(kind), samples=4, seed=5)
    assert st.queries == nq
    # a v_scale below 1 takes asin past 1 at the sphere's poles: NaN there,
    # in the reference as here
    both_nan = np.isnan(img) & np.isnan(ref)
    exact = ((img.view(np.uint32) == ref.view(np.uint32)) | both_nan).all(axis=2)
    close = np.isclose(img, ref, rtol=1e-6, atol=1e-7, equal_nan=True)
    assert close.all(), f"{(~close).sum()} values differ"
    assert exact.mean() >= 0.995, f"only {exact.mean():.4f} of pixels bit-exact"
    assert len(np.unique(ref[..., 0])) > 100


@pytest.mark.gpu
def test_xtrender_exports(tmp_path):
    """The headless driver's export: -outdir names the file as the reference
    CLI does (<scene>_<renderer>.png) and its pixels are the 8-bit conversion
    of the same render written raw."""
    import os
    import shutil
    import subprocess
    exe = os.path.join(os.path.dirname(xa.__file__), "xtrender")
    scn = tmp_path / "spheres.scn"
    shutil.copy(scene_path("spheres"), scn)
    args = [exe, str(scn), "-renderer", "stencil", "-res", "40x24", "-samples", "2"]
    for extra in (["-out", str(tmp_path / "frame.raw")], ["-outdir", str(tmp_path / "out")]):
        os.makedirs(tmp_path / "out", exist_ok=True)
        r = subprocess.run(args + extra, capture_output=True, text=True, timeout=120)
        assert r.returncode == 0, r.stderr
    raw = np.fromfile(tmp_path / "frame.raw", np.float32).reshape(24, 40, 4)
    from tests.test_image_io import read_png, to8
    png = read_png((tmp_path / "out" / "spheres_stencil.png").read_bytes())
    assert np.array_equal(png, to8(raw))
    assert raw[..., 3].max() > 0


@pytest.mark.gpu
@pytest.mark.parametrize("integrator", [xa.STENCIL, xa.DEPTH, xa.NORMAL, xa.EMISSION])
@pytest.mark.parametrize("kind", ["erp", "ods", "cubemap"])
@pytest.mark.parametrize("accel", [xa.ACCEL_KDTREE, xa.ACCEL_OCTREE])
def test_camera_kinds(device, kind, integrator, accel, tmp_path):
    """ODS / ERP / cubemap primary rays, evaluated in FP64 on the device with
    the reference's float/double mix. The cubemap has no transcendental
    functions: its images are bit-exact against the oracle. ERP and ODS take
    double sines and cosines, which the device's math library and the host's
    may round differently in the last bit: pixels must agree to 1e-5
    relative, and all but 0.2% of them bit for bit."""
    from tests.helpers import camera_dict_of, camera_scene
    s = xa.Scene(camera_scene(str(tmp_path / "cam.scn"), kind))
    sid = device.upload(s)
    W, H = 96, 72
    job = xa.make_job(sid, integrator, W, H, s.camera(), samples=4, seed=5, accel=accel)
    img, st = device.render(job)
    ref, nq = OracleScene(s).render(integrator, W, H, camera_dict_of(kind), samples=4, seed=5)
    assert st.queries == nq
    exact = (img.view(np.uint32) == ref.view(np.uint32)).all(axis=2)
    close = np.isclose(img, ref, rtol=1e-5, atol=1e-6).all(axis=2)
    if kind == "cubemap":
        assert exact.all(), f"{(~exact).sum()} pixels differ"
    else:
        assert close.all(), f"{(~close).sum()} pixels differ by more than 1e-5"
        assert exact.mean() >= 0.998, f"only {exact.mean():.4f} of pixels bit-exact"
    assert (ref[..., 3] > 0).sum() > 1000
    device.release(sid)


@pytest.mark.gpu
def test_cubemap_height_refused(device, tmp_path):
    from tests.helpers import camera_scene
    s = xa.Scene(camera_scene(str(tmp_path / "cam.scn"), "cubemap"))
    sid = device.upload(s)
    with pytest.raises(xa.XtError) as e:
        device.render(xa.make_job(sid, xa.STENCIL, 32, 40, s.camera()))
    assert e.value.code == 1
    device.release(sid)


@pytest.mark.gpu
@pytest.mark.parametrize("kind", ["pinhole", "erp"])
def test_dielectric_same_stream(device, kind, tmp_path):
    """A dielectric sphere in view ends its paths with 0, as the reference's
    zero intensity does; with the same seed device and oracle agree on all
    but the FP32-decided pixels, and on the query count."""
    from tests.helpers import camera_dict_of, camera_scene
    s = xa.Scene(camera_scene(str(tmp_path / "cam.scn"), kind))
    sid = device.upload(s)
    W, H, spp = 64, 48, 8
    job = xa.make_job(sid, xa.PATHTRACER, W, H, s.camera(), samples=spp, rdepth=6, seed=11,
                      accum=xa.ACCUM_REFERENCE)
    img, st = device.render(job)
    ref, nq, mean, m2 = OracleScene(s).render(xa.PATHTRACER, W, H, camera_dict_of(kind), samples=spp, rdepth=6, seed=11,
                                              stats=True)
    same_stream_check(img, ref, mean, m2, spp, 0.98)
    assert abs(st.queries - nq) <= 0.01 * nq
    device.release(sid)


@pytest.mark.gpu
def test_raygraph_equals_reference_file(device, tmp_path):
    """The ray graph of a depth render (33x33, GRID aa 2: partial 32-pixel
    tiles, four samples per pixel) is byte-identical to the PLY the
    reference's raygraph::write produced for the same render
    (tests/golden/ref_raygraph_spheres.npz), and the image is too."""
    import os
    z = np.load(os.path.join(os.path.dirname(__file__), "golden", "ref_raygraph_spheres.npz"))
    s = xa.Scene(scenes.SCENES["spheres"][0]())
    sid = device.upload(s)
    job = xa.make_job(sid, xa.DEPTH, int(z["width"]), int(z["height"]), s.camera(), samples=int(z["samples"]),
                      aa=int(z["aa"]))
    img = device.render_raygraph(job, tmp_path / "rg.ply")
    assert np.array_equal(img.view(np.uint32), z["image"].view(np.uint32))
    got = np.fromfile(tmp_path / "rg.ply", np.uint8)
    ref = z["ply"]
    if not np.array_equal(got, ref):
        a, b = got.tobytes().decode().splitlines(), ref.tobytes().decode().splitlines()
        diff = [(k, x, y) for k, (x, y) in enumerate(zip(a, b)) if x != y][:5]
        raise AssertionError(f"{len(a)} vs {len(b)} lines; first differences {diff}")
    # other integrators record no paths: a header with empty elements
    device.render_raygraph(xa.make_job(sid, xa.STENCIL, 8, 8, s.camera()), tmp_path / "st.ply")
    text = (tmp_path / "st.ply").read_text()
    assert "element vertex 0\n" in text and "element edge 0\n" in text and text.endswith("end_header\n")
    device.release(sid)


@pytest.mark.gpu
def test_ao_and_depth_on_the_octree_at_256(device):
    """The CPU-runnable configuration: ground plane and 64 Lambert spheres,
    pinhole camera, 256x256, the reference's octree; ao with 16 hemisphere
    samples per pixel on the oracle's random streams, depth bit-exact."""
    s, sid, cam = _setup(device, "spheres")
    W = H = 256
    job = xa.make_job(sid, xa.AO, W, H, s.camera(), samples=16, seed=5, accel=xa.ACCEL_OCTREE,
                      accum=xa.ACCUM_REFERENCE)
    img, st = device.render(job)
    ref, nq, mean, m2 = OracleScene(s).render(xa.AO, W, H, cam, samples=16, seed=5, stats=True)
    same_stream_check(img, ref, mean, m2, 16, 0.995)
    assert abs(st.queries - nq) <= 0.005 * nq
    job = xa.make_job(sid, xa.DEPTH, W, H, s.camera(), samples=1, accel=xa.ACCEL_OCTREE)
    img, st = device.render(job)
    ref, nq = OracleScene(s).render(xa.DEPTH, W, H, cam, samples=1)
    assert st.queries == nq and np.array_equal(img.view(np.uint32), ref.view(np.uint32))
    device.release(sid)
