"""Continuous-batching server: served images equal one-job renders bit for
bit, jobs join at pass boundaries, admission follows the HBM budget."""
import numpy as np
import pytest

import xtracer_amd as xa
from tests.helpers import scene_path


@pytest.fixture
def small_passes(device):
    device.set_tuning(wave_paths=1 << 16)
    yield device
    device.set_tuning(wave_paths=16 << 20)


def _scenes(device):
    out = {}
    for kind, kw in (("spheres", {}), ("icosphere", {"subdiv": 3})):
        s = xa.Scene(scene_path(kind, **kw))
        out[kind] = (s, device.upload(s))
    return out


def _drain_all(device, tickets):
    device.drain()
    got = {}
    for t, (w, h) in tickets.items():
        done, img, st = device.poll(t, w, h)
        assert done
        got[t] = (img, st)
    return got


@pytest.mark.gpu
@pytest.mark.parametrize("policy", ["srpt", "fair"])
def test_served_equals_alone(small_passes, policy):
    dev = small_passes
    dev.set_serve_policy(policy)
    sc = _scenes(dev)
    jobs = []
    for k, (kind, integ, w, h, spp) in enumerate([
            ("spheres", xa.STENCIL, 64, 48, 4), ("spheres", xa.DEPTH, 80, 40, 2),
            ("icosphere", xa.AO, 72, 56, 8), ("icosphere", xa.PATHTRACER, 96, 64, 16),
            ("spheres", xa.PATHTRACER, 40, 40, 5), ("icosphere", xa.PATHTRACER, 33, 17, 3),
            ("icosphere", xa.PATHTRACER, 320, 240, 3),  # more tile positions than a pass: sliced samples
            ("icosphere", xa.NORMAL, 64, 64, 2), ("spheres", xa.EMISSION, 48, 32, 4),
            ("icosphere", xa.WIREFRAME, 40, 24, 3)]):
        s, sid = sc[kind]
        jobs.append(xa.make_job(sid, integ, w, h, s.camera(), samples=spp, rdepth=5, seed=7 + k, job_key=k))
    tickets = {dev.submit(j): (j.width, j.height) for j in jobs}
    got = _drain_all(dev, tickets)
    assert dev.serve_info()["jobs_active"] == 0
    dev.set_serve_policy("srpt")
    for (t, (img, st)), j in zip(got.items(), jobs):
        ref, _ = dev.render(j)
        assert np.array_equal(img.view(np.uint32), ref.view(np.uint32)), f"job {t} differs from its one-job render"
        assert st.latency_ms >= st.queue_ms >= 0.0
        assert st.samples == j.samples * j.width * j.height


@pytest.mark.gpu
def test_small_job_overtakes_large(small_passes):
    dev = small_passes
    s = xa.Scene(scene_path("icosphere", subdiv=3))
    sid = dev.upload(s)
    big = xa.make_job(sid, xa.PATHTRACER, 256, 256, s.camera(), samples=32, rdepth=4, job_key=1)
    small = xa.make_job(sid, xa.PATHTRACER, 32, 32, s.camera(), samples=2, rdepth=4, job_key=2)
    tb = dev.submit(big)
    dev.serve_step()
    ts = dev.submit(small)
    steps = 0
    while True:
        steps += 1
        dev.serve_step()
        done, img, _ = dev.poll(ts, 32, 32)
        if done:
            break
        assert steps < 50
    done_big, _, st = dev.poll(tb, 256, 256)
    assert not done_big and st.passes >= 1
    dev.drain()
    done_big, img_big, _ = dev.poll(tb, 256, 256)
    assert done_big
    ref, _ = dev.render(big)
    assert np.array_equal(img_big.view(np.uint32), ref.view(np.uint32))


@pytest.mark.gpu
def test_budget_admission():
    s = xa.Scene(scene_path("spheres"))
    dev = xa.Device(0, hbm_budget=512 << 20)
    try:
        dev.set_tuning(wave_paths=1 << 16)
        sid = dev.upload(s)
        with pytest.raises(xa.XtError) as e:
            dev.submit(xa.make_job(sid, xa.STENCIL, 4096, 4096, s.camera()))  # 1.3 GB of job buffers
        assert e.value.code == 5
        # two jobs of 320 MB each: the second waits for the first's buffers
        w, h = 2000, 2000
        j1 = xa.make_job(sid, xa.STENCIL, w, h, s.camera(), job_key=1)
        j2 = xa.make_job(sid, xa.STENCIL, w, h, s.camera(), job_key=2)
        t1, t2 = dev.submit(j1), dev.submit(j2)
        dev.serve_step()
        info = dev.serve_info()
        assert info["jobs_active"] == 1 and info["jobs_waiting"] == 1
        dev.drain()
        d1, i1, s1 = dev.poll(t1, w, h)
        d2, i2, s2 = dev.poll(t2, w, h)
        assert d1 and d2
        assert s2.queue_ms >= s1.queue_ms
        assert np.array_equal(i1.view(np.uint32), i2.view(np.uint32))
        with pytest.raises(xa.XtError):
            dev.poll(t1, w, h)  # a finished ticket is forgotten
    finally:
        dev.close()


@pytest.mark.gpu
def test_scene_in_use(device):
    s = xa.Scene(scene_path("spheres"))
    sid = device.upload(s)
    t = device.submit(xa.make_job(sid, xa.STENCIL, 32, 32, s.camera()))
    with pytest.raises(xa.XtError) as e:
        device.release(sid)
    assert e.value.code == 7
    device.drain()
    assert device.poll(t, 32, 32)[0]
    device.release(sid)
