"""Launch-shape knobs change speed, never results: every kd stack depth, trace
block size, persistent grid size and wavefront size gives the same hits and
the same bit-exact images as the defaults."""
import numpy as np
import pytest

import xtracer_amd as xa
from oracle.oracle import OracleScene
from xtracer_amd import scenes
from tests.helpers import mixed_rays, scene_path

DEFAULT = {"kd_stack": 12, "trace_block": 256, "wave_paths": 64 << 20, "trace_blocks_per_cu": xa.TUNING_AUTO,
           "trace_segments": 1, "ray_order": 1, "kd_layout": 1, "kd_leaf": 1, "frame_lanes": 2}


@pytest.fixture
def tuned(device):
    yield device
    device.set_tuning(**DEFAULT)


@pytest.mark.gpu
def test_tuning_roundtrip_and_rejects(tuned):
    t = tuned.set_tuning(**DEFAULT)
    assert t == dict(DEFAULT, trace_blocks_per_cu=0)
    for bad in ({"kd_stack": 10}, {"trace_block": 64}, {"wave_paths": 100}, {"trace_blocks_per_cu": 65},
                {"trace_segments": 4}, {"ray_order": 4}, {"kd_layout": 3}, {"kd_leaf": 3},
                {"frame_lanes": 5}):
        with pytest.raises(xa.XtError):
            tuned.set_tuning(**bad)
    assert tuned.tuning() == t


@pytest.mark.gpu
@pytest.mark.parametrize("stack", [8, 12, 16, 24])
@pytest.mark.parametrize("block", [128, 256])
@pytest.mark.parametrize("per_cu", [1, xa.TUNING_AUTO])
def test_trace_invariant(tuned, stack, block, per_cu):
    s = xa.Scene(scene_path("icosphere", subdiv=5))
    sid = tuned.upload(s)
    o, d = mixed_rays(None, 1 << 16, seed=5)
    tuned.set_tuning(**DEFAULT)
    base = tuned.trace(sid, o, d, xa.ACCEL_KDTREE)
    tuned.set_tuning(kd_stack=stack, trace_block=block, trace_blocks_per_cu=per_cu)
    got = tuned.trace(sid, o, d, xa.ACCEL_KDTREE)
    tuned.release(sid)
    assert np.array_equal(base[1], got[1])
    # a ray may take the FP64 re-test under one stack depth and not another
    np.testing.assert_allclose(got[0], base[0], rtol=1e-6)


@pytest.mark.gpu
@pytest.mark.parametrize("integrator", [xa.STENCIL, xa.DEPTH])
def test_render_bit_exact_small_shapes(tuned, integrator):
    """Several wavefront passes, one resident block per CU and an 8-entry
    stack still reproduce the oracle's image bit for bit."""
    s = xa.Scene(scene_path("icosphere", subdiv=4))
    sid = tuned.upload(s)
    cam = scenes.camera_dict(**scenes.SCENES["icosphere"][1])
    W, H, spp = 96, 64, 4
    tuned.set_tuning(kd_stack=8, trace_block=128, trace_blocks_per_cu=1, wave_paths=4096)
    img, st = tuned.render(xa.make_job(sid, integrator, W, H, s.camera(), samples=spp, seed=5))
    tuned.release(sid)
    assert st.passes == spp  # 4096 paths per pass: one sample per pass
    ref, nq = OracleScene(s).render(integrator, W, H, cam, samples=spp, seed=5)
    assert st.queries == nq
    assert np.array_equal(img.view(np.uint32), ref.view(np.uint32))


LAYOUTS = [{"trace_segments": 8}, {"ray_order": 2}, {"ray_order": 3}, {"kd_layout": 2}, {"kd_leaf": 2},
           {"frame_lanes": 1}, {"frame_lanes": 4},
           {"trace_segments": 8, "ray_order": 2, "kd_layout": 2, "kd_leaf": 2, "frame_lanes": 2}]


@pytest.mark.gpu
@pytest.mark.parametrize("knobs", LAYOUTS)
def test_layout_knobs_invariant(tuned, knobs):
    """Per-XCD queue segments, tile-major and pixel-major camera rays, the
    128-byte-line node order and leaves read through the reference list give
    the default's hits and the oracle's bit-exact depth image."""
    s = xa.Scene(scene_path("icosphere", subdiv=5))
    tuned.set_tuning(**DEFAULT)
    sid = tuned.upload(s)
    o, d = mixed_rays(None, 1 << 16, seed=9)
    base = tuned.trace(sid, o, d, xa.ACCEL_KDTREE)
    tuned.set_tuning(**knobs)
    sid2 = tuned.upload(s)  # kd_layout applies to scenes uploaded afterwards
    got = tuned.trace(sid2, o, d, xa.ACCEL_KDTREE)
    assert np.array_equal(base[1], got[1])
    np.testing.assert_allclose(got[0], base[0], rtol=1e-6)
    cam = scenes.camera_dict(**scenes.SCENES["icosphere"][1])
    W, H, spp = 88, 56, 3
    img, st = tuned.render(xa.make_job(sid2, xa.DEPTH, W, H, s.camera(), samples=spp, seed=4))
    tuned.release(sid)
    tuned.release(sid2)
    ref, nq = OracleScene(s).render(xa.DEPTH, W, H, cam, samples=spp, seed=4)
    assert st.queries == nq
    assert np.array_equal(img.view(np.uint32), ref.view(np.uint32))


@pytest.mark.gpu
@pytest.mark.parametrize("integrator", [xa.PATHTRACER, xa.AO, xa.NORMAL, xa.DEPTH])
@pytest.mark.parametrize("ray_order", [1, 2, 3])
@pytest.mark.parametrize("lanes", [2, 4])
def test_frame_lanes_same_bits(tuned, integrator, ray_order, lanes):
    """A one-job render whose passes run as 2 or 4 pixel groups on streams of
    their own gives the one-group image bit for bit and the same query count,
    over several passes, with an odd number of 8x8 tiles (77: groups of 38 and
    39, or 19, 19, 19 and 20 tiles), and with paths that all end long before
    the depth limit of 100 (the early stop)."""
    s = xa.Scene(scene_path("spheres"))
    sid = tuned.upload(s)
    W, H, spp = 88, 56, 6
    job = xa.make_job(sid, integrator, W, H, s.camera(), samples=spp, rdepth=100, seed=21)
    tuned.set_tuning(**dict(DEFAULT, wave_paths=3 * 77 * 64, ray_order=ray_order, frame_lanes=1))
    ref, st1 = tuned.render(job)
    tuned.set_tuning(frame_lanes=lanes)
    img, st2 = tuned.render(job)
    tuned.release(sid)
    assert st1.passes == st2.passes == 2
    assert st2.queries == st1.queries
    assert np.array_equal(img.view(np.uint32), ref.view(np.uint32))


@pytest.mark.gpu
def test_trace_busy_time(tuned):
    """The job stats' trace busy time is the union of the timed trace
    launches' intervals: with one group per pass the launches do not overlap
    and it equals the sum of their times; with two groups it lies between
    the longest launch and that sum, and within the render's time."""
    s = xa.Scene(scene_path("icosphere", subdiv=5))
    sid = tuned.upload(s)
    job = xa.make_job(sid, xa.PATHTRACER, 320, 240, s.camera(), samples=8, rdepth=6, seed=2)
    tuned.set_tuning(**dict(DEFAULT, frame_lanes=1))
    _, one = tuned.render(job)
    tuned.set_tuning(frame_lanes=2)
    _, two = tuned.render(job)
    tuned.release(sid)
    assert one.trace_launches > 0 and two.trace_launches > one.trace_launches
    assert one.ms_trace_busy == pytest.approx(one.ms_trace, rel=1e-3, abs=1e-3)
    assert 0.0 < two.ms_trace_busy <= two.ms_trace * (1 + 1e-4) + 1e-3
    assert two.ms_trace_busy <= two.ms_total * (1 + 1e-4) + 1e-3
    assert two.ms_trace_busy >= two.ms_trace / two.trace_launches


@pytest.mark.gpu
def test_autotune_stages(tuned):
    """The staged search tries every value of every stage, leaves the device on
    the fastest shape it timed, and that shape renders the same bits."""
    from xtracer_amd.autotune import autotune
    from tests.helpers import HipBuffers
    s = xa.Scene(scene_path("icosphere", subdiv=4))
    sid = tuned.upload(s)
    W, H = 64, 48
    hip = HipBuffers()
    out = hip.put(np.zeros((H, W, 4), np.float32))
    try:
        job = xa.make_job(sid, xa.PATHTRACER, W, H, s.camera(), samples=4, rdepth=4, seed=3)
        ref, _ = tuned.render(job)
        stages = [{"kd_stack": (8, 16)}, {"ray_order": (1, 2, 3)}, {"trace_segments": (1, 8)}]
        best, table = autotune(tuned, job, out.value, samples=2, repeats=1, wave_paths=(1 << 16,), stages=stages)
        assert len(table) == 2 + 3 + 2 + 2
        assert {t["kd_stack"] for t, _ in table[:2]} == {8, 16}
        assert {t["ray_order"] for t, _ in table[2:5]} == {1, 2, 3}
        assert tuned.tuning() == best
        assert best["kd_stack"] in (8, 16) and best["ray_order"] in (1, 2, 3) and best["trace_segments"] in (1, 8)
        img, _ = tuned.render(job)
        assert np.array_equal(img.view(np.uint32), ref.view(np.uint32))
    finally:
        hip.free()
        tuned.release(sid)
