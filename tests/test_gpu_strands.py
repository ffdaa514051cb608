"""The incoherent-ray stress workload: random-walk strands of thin triangles
in the unit sphere, one mesh light, Blinn-Phong, pathtracer with max depth 6,
kd-tree. At a size the oracle renders in seconds the device traces the
oracle's paths (same counter-based streams) and agrees with an independent
oracle estimate within the Monte Carlo bound; at the full 50,000 strands
(2.95 M triangles, 1024x1024, 32 spp) the render is checked through
properties: hits against the oracle on a ray sample, a finite non-negative
image, the same bits on a second run and on another launch shape."""
import numpy as np
import pytest

import xtracer_amd as xa
from oracle.oracle import OracleScene
from xtracer_amd import scenes
from tests.helpers import same_stream_check, soup_to_object_local

CAM = scenes.camera_dict(**scenes.STRANDS_CAMERA)


@pytest.mark.gpu
def test_strands_pathtracer_same_stream(device):
    s = xa.Scene(scenes.strands_scene(strands=600))
    sid = device.upload(s)
    W, H, spp = 64, 64, 8
    job = xa.make_job(sid, xa.PATHTRACER, W, H, s.camera(), samples=spp, rdepth=6, seed=21,
                      accum=xa.ACCUM_REFERENCE)
    img, st = device.render(job)
    ref, nq, mean, m2 = OracleScene(s).render(xa.PATHTRACER, W, H, CAM, samples=spp, rdepth=6, seed=21, stats=True)
    same_stream_check(img, ref, mean, m2, spp, 0.995)
    assert abs(st.queries - nq) <= 0.01 * nq
    assert ref[..., :3].max() > 0.0  # the light and the Blinn-Phong lobes are seen
    device.release(sid)


@pytest.mark.gpu
def test_strands_pathtracer_monte_carlo_bound(device):
    s = xa.Scene(scenes.strands_scene(strands=600))
    sid = device.upload(s)
    W, H, spp = 24, 24, 64
    img, _ = device.render(xa.make_job(sid, xa.PATHTRACER, W, H, s.camera(), samples=spp, rdepth=6, seed=101))
    _, _, mean, m2 = OracleScene(s).render(xa.PATHTRACER, W, H, CAM, samples=4 * spp, rdepth=6, seed=202, stats=True)
    var = m2 / max(1, 4 * spp - 1)
    se = np.sqrt(var / spp + var / (4 * spp))
    inside = np.abs(img[..., :3] - mean) <= 2.576 * se + 1e-6
    assert inside.mean() >= 0.97, f"{inside.mean():.4f} of channels inside the 99% bound"
    device.release(sid)


@pytest.mark.gpu
def test_strands_full_size_properties(device):
    s = xa.Scene(scenes.strands_scene())
    info = s.info()
    assert info.n_prims >= 2_950_000
    sid = device.upload(s)
    # hits of camera rays against the oracle's FP64 restatement on a sample
    rng = np.random.default_rng(4)
    xy = np.stack([rng.uniform(0, 1024, 1024), rng.uniform(0, 1024, 1024)], 1).astype(np.float32)
    from oracle.oracle import camera_rays
    o, d = camera_rays(CAM, 1024, 1024, xy)
    o32, d32 = o.astype(np.float32), d.astype(np.float32)
    t, prim, _, _, _ = device.trace(sid, o32, d32, xa.ACCEL_KDTREE)
    tr, obr, pr, _, _ = OracleScene(s).trace(o32.astype(np.float64), d32.astype(np.float64))
    obj, loc = soup_to_object_local(s.arrays(), prim)
    same = (obj == obr) & ((obr < 0) | (loc == pr))
    assert same.mean() >= 0.999, f"{(~same).sum()} of {len(same)} camera rays disagree"
    assert (obr >= 0).mean() > 0.1  # the strands fill part of the view
    # the full frame: 1024x1024, 32 spp, max depth 6
    job = xa.make_job(sid, xa.PATHTRACER, 1024, 1024, s.camera(), samples=32, rdepth=6, seed=9)
    a, st = device.render(job)
    assert np.isfinite(a).all() and (a[..., :3] >= 0).all()
    assert st.queries >= 1024 * 1024 * 32 and st.trace_fail == 0
    assert a[..., :3].mean() > 0.0
    b, st2 = device.render(job)
    assert np.array_equal(a.view(np.uint32), b.view(np.uint32)) and st2.queries == st.queries
    t0 = device.tuning()
    try:
        device.set_tuning(kd_stack=24, wave_paths=1 << 22)
        c, st3 = device.render(job)
    finally:
        device.set_tuning(**{k: v for k, v in t0.items() if k != "trace_blocks_per_cu"},
                          trace_blocks_per_cu=t0["trace_blocks_per_cu"] or xa.TUNING_AUTO)
    assert np.array_equal(a.view(np.uint32), c.view(np.uint32)) and st3.queries == st.queries
    device.release(sid)
