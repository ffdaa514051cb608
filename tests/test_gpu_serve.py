"""Continuous-batching server.

Every served job equals its one-job render bit for bit: jobs in whole passes
(the deterministic integrators, and every job with the regeneration engine
off) accumulate like xt_render, and ao / pathtracer jobs on the
path-regeneration engine keep each sample's radiance and sum it per pixel in
sample order when they finish. Jobs too large for a sample buffer sum in
64-bit fixed point: their images do not depend on the schedule (graph or
launch by launch, policy, mix, path slots) and differ from the one-job
render's compensated FP64 sum only by the rounding of the two sums. Jobs join
between rounds; admission follows the HBM budget."""
import numpy as np
import pytest

import xtracer_amd as xa
from oracle.oracle import OracleScene
from xtracer_amd import scenes
from tests.helpers import same_stream_check, scene_path

# served (fixed-point sum) against xt_render (Neumaier FP64 sum): both are
# exact to far below float resolution, so they agree to a few float ulps
REGEN_RTOL, REGEN_ATOL = 1e-6, 1e-9


@pytest.fixture
def small_passes(device):
    """Passes and the regeneration engine's slot pool of 64 Ki paths, so that
    larger jobs take many passes / rounds."""
    device.drain()
    device.set_tuning(wave_paths=1 << 16)
    device.serve_configure(regen=True, graph=True, rounds_per_step=4, paths=1 << 16)
    yield device
    device.drain()
    device.set_tuning(wave_paths=64 << 20)
    device.set_serve_policy("srpt")
    device.serve_configure(regen=True, graph=True, rounds_per_step=4, paths=16 << 20, loop=True)


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


def _regen(job):
    return job.integrator in (xa.AO, xa.PATHTRACER) and job.accum == xa.ACCUM_DEFAULT


def _same(img, ref, fixed=False):
    if not fixed:
        return np.array_equal(img.view(np.uint32), ref.view(np.uint32))
    return np.allclose(img, ref, rtol=REGEN_RTOL, atol=REGEN_ATOL)


MIX = [("spheres", xa.STENCIL, 64, 48, 4), ("spheres", xa.DEPTH, 80, 40, 2),
       ("icosphere", xa.AO, 72, 56, 8), ("icosphere", xa.PATHTRACER, 96, 64, 16),
       ("spheres", xa.PATHTRACER, 40, 40, 5), ("icosphere", xa.PATHTRACER, 33, 17, 3),
       ("icosphere", xa.PATHTRACER, 320, 240, 3),  # more tile positions than a pass: sliced samples
       ("icosphere", xa.NORMAL, 64, 64, 2), ("spheres", xa.EMISSION, 48, 32, 4),
       ("icosphere", xa.WIREFRAME, 40, 24, 3), ("spheres", xa.AO, 56, 40, 6)]


def _mix_jobs(sc):
    jobs = []
    for k, (kind, integ, w, h, spp) in enumerate(MIX):
        s, sid = sc[kind]
        jobs.append(xa.make_job(sid, integ, w, h, s.camera(), samples=spp, rdepth=5, seed=7 + k, job_key=k))
    return jobs


@pytest.mark.gpu
@pytest.mark.parametrize("policy", ["srpt", "fair", "fcfs", "hybrid"])
@pytest.mark.parametrize("regen,loop", [(True, True), (True, False), (False, True)])
def test_served_equals_alone(small_passes, policy, regen, loop):
    """loop: device-fed (True) or host-stepped (False) regeneration engine."""
    dev = small_passes
    dev.set_serve_policy(policy)
    dev.serve_configure(regen=regen, loop=loop)
    sc = _scenes(dev)
    jobs = _mix_jobs(sc)
    tickets = {dev.submit(j): (j.width, j.height) for j in jobs}
    got = _drain_all(dev, tickets)
    assert dev.serve_info()["jobs_active"] == 0
    for (t, (img, st)), j in zip(got.items(), jobs):
        ref, _ = dev.render(j)
        assert _same(img, ref), f"job {t} ({j.integrator}) differs from its one-job render"
        assert st.latency_ms >= st.queue_ms >= 0.0
        assert st.samples == j.samples * j.width * j.height


@pytest.mark.gpu
def test_regen_images_independent_of_schedule(small_passes):
    """Graph replay or kernel-by-kernel launches, a round's lanes concurrent or
    one after another, every policy, other path-slot counts and rounds per
    step, alone or in the mix: bit-identical images."""
    dev = small_passes
    sc = _scenes(dev)
    jobs = [j for j in _mix_jobs(sc) if _regen(j)]
    runs = []
    try:
        for policy, graph, rounds, par in (("srpt", True, 4, True), ("fcfs", False, 2, True), ("hybrid", True, 8, False),
                                           ("fair", False, 6, False)):
            dev.set_serve_policy(policy)
            dev.serve_configure(graph=graph, rounds_per_step=rounds, lane_streams=par)
            tickets = {dev.submit(j): (j.width, j.height) for j in jobs}
            got = _drain_all(dev, tickets)
            runs.append([img for img, _ in got.values()])
    finally:
        dev.serve_configure(lane_streams=True)
    for r in runs[1:]:
        for a, b in zip(runs[0], r):
            assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
    for k, j in enumerate(jobs):  # one job at a time
        t = dev.submit(j)
        alone = _drain_all(dev, {t: (j.width, j.height)})[t][0]
        assert np.array_equal(alone.view(np.uint32), runs[0][k].view(np.uint32)), f"job {k} alone"


@pytest.mark.gpu
def test_regen_fixed_point_jobs(small_passes):
    """With a 1 MiB sample-buffer limit the larger jobs sum in fixed point:
    within the rounding of the two sums of their one-job renders, and the same
    bits under every schedule."""
    dev = small_passes
    dev.serve_configure(sample_buffer_mb=1)
    try:
        sc = _scenes(dev)
        jobs = [j for j in _mix_jobs(sc) if _regen(j)]
        fixed = [j.width * j.height * j.samples * 16 > (1 << 20) for j in jobs]
        assert any(fixed) and not all(fixed)
        runs = []
        for policy, graph, par in (("srpt", True, True), ("fcfs", False, False)):
            dev.set_serve_policy(policy)
            dev.serve_configure(graph=graph, lane_streams=par)
            tickets = {dev.submit(j): (j.width, j.height) for j in jobs}
            runs.append([img for img, _ in _drain_all(dev, tickets).values()])
        for a, b in zip(*runs):
            assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
        for img, j, fx in zip(runs[0], jobs, fixed):
            ref, _ = dev.render(j)
            assert _same(img, ref, fx)
    finally:
        dev.serve_configure(sample_buffer_mb=4096, graph=True, lane_streams=True)


@pytest.mark.gpu
def test_regen_other_path_slot_count(device):
    """A smaller slot pool (more regeneration per path) gives the same images."""
    device.set_tuning(wave_paths=1 << 16)
    try:
        sc = _scenes(device)
        jobs = [j for j in _mix_jobs(sc) if _regen(j)][:3]
        imgs = []
        for paths in (1 << 16, 1 << 13):
            device.drain()
            device.serve_configure(regen=True, paths=paths)
            assert device.serve_config()["paths"] == paths
            tickets = {device.submit(j): (j.width, j.height) for j in jobs}
            imgs.append([img for img, _ in _drain_all(device, tickets).values()])
        for a, b in zip(*imgs):
            assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
    finally:
        device.serve_configure(paths=16 << 20)
        device.set_tuning(wave_paths=64 << 20)


@pytest.mark.gpu
def test_regen_same_stream_as_oracle(small_passes):
    """A served pathtracer job traces the oracle's paths (same counter-based
    streams): pixels agree except where an FP32 decision differs."""
    dev = small_passes
    s = xa.Scene(scene_path("icosphere", subdiv=4))
    sid = dev.upload(s)
    cam = scenes.camera_dict(**scenes.SCENES["icosphere"][1])
    W, H, spp = 64, 48, 8
    job = xa.make_job(sid, xa.PATHTRACER, W, H, s.camera(), samples=spp, rdepth=6, seed=11)
    t = dev.submit(job)
    img = _drain_all(dev, {t: (W, H)})[t][0]
    ref, _, mean, m2 = OracleScene(s).render(xa.PATHTRACER, W, H, cam, samples=spp, rdepth=6, seed=11, stats=True)
    same_stream_check(img, ref, mean, m2, spp, 0.995)


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
    assert _same(img_big, ref)
    ref_small, _ = dev.render(small)
    assert _same(img, ref_small)


@pytest.mark.gpu
@pytest.mark.parametrize("regen", [True, False])
def test_budget_admission(regen):
    s = xa.Scene(scene_path("spheres"))
    dev = xa.Device(0, hbm_budget=512 << 20)
    try:
        dev.set_tuning(wave_paths=1 << 16)
        dev.serve_configure(regen=regen, paths=1 << 16)
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
        # ao jobs of 160 MB on the regeneration engine (40 B per pixel): the
        # second waits likewise
        j3 = xa.make_job(sid, xa.AO, w, h, s.camera(), samples=1, job_key=3)
        j4 = xa.make_job(sid, xa.AO, w, h, s.camera(), samples=1, job_key=4)
        t3, t4 = dev.submit(j3), dev.submit(j4)
        dev.drain()
        assert dev.poll(t3, w, h)[0] and dev.poll(t4, w, h)[0]
    finally:
        dev.close()


@pytest.mark.gpu
def test_serve_config_rejects(device):
    with pytest.raises(xa.XtError):
        device.serve_configure(rounds_per_step=3)
    with pytest.raises(xa.XtError):
        device.serve_configure(paths=100)
    with pytest.raises(xa.XtError):
        device.set_serve_policy(7)


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
    t = device.submit(xa.make_job(sid, xa.AO, 32, 32, s.camera(), samples=2))
    with pytest.raises(xa.XtError) as e:
        device.release(sid)
    device.drain()
    assert device.poll(t, 32, 32)[0]
    device.release(sid)


@pytest.mark.gpu
@pytest.mark.parametrize("policy", ["srpt", "fair", "fcfs", "hybrid"])
@pytest.mark.parametrize("graph", [False])
def test_device_fed_equals_alone(small_passes, policy, graph):
    """Device-fed serving (jobs reach the device through a request ring in
    host memory, the device orders them by the policy, resolves finished
    ones and posts them to a completion ring): every job of the mix equals
    its one-job render bit for bit, the deterministic integrators running in
    whole passes beside them."""
    dev = small_passes
    dev.set_serve_policy(policy)
    dev.serve_configure(loop=True, graph=graph)
    try:
        assert dev.serve_config()["loop"] == 2
        sc = _scenes(dev)
        jobs = _mix_jobs(sc)
        tickets = {dev.submit(j): (j.width, j.height) for j in jobs}
        got = _drain_all(dev, tickets)
        assert dev.serve_info()["jobs_active"] == 0
        for (t, (img, st)), j in zip(got.items(), jobs):
            ref, _ = dev.render(j)
            assert _same(img, ref), f"job {t} ({j.integrator}) differs from its one-job render"
            assert st.latency_ms >= st.queue_ms >= 0.0
            assert st.samples == j.samples * j.width * j.height
    finally:
        dev.serve_configure(loop=True, graph=True)


@pytest.mark.gpu
def test_device_fed_arrivals_and_fixed_point(small_passes):
    """Jobs submitted one per step while earlier ones are on the device, the
    larger ones over a 1 MiB sample-buffer limit (fixed-point sums), three
    rounds of the mix so that table entries are reused: every image is the
    host-stepped engine's bit for bit."""
    dev = small_passes
    dev.serve_configure(sample_buffer_mb=1, loop=False)
    try:
        sc = _scenes(dev)
        base = [j for j in _mix_jobs(sc) if _regen(j)]
        jobs = [type(j).from_buffer_copy(j) for _ in range(3) for j in base]
        for k, j in enumerate(jobs):
            j.job_key = 100 + k
        tickets = {dev.submit(j): (j.width, j.height) for j in jobs}
        ref = [img for img, _ in _drain_all(dev, tickets).values()]
        dev.serve_configure(loop=True)
        pending, got, k, steps = {}, {}, 0, 0
        while k < len(jobs) or pending:
            if k < len(jobs):
                pending[dev.submit(jobs[k])] = k
                k += 1
            dev.serve_step()
            steps += 1
            for t in list(pending):
                j = jobs[pending[t]]
                done, img, _ = dev.poll(t, j.width, j.height)
                if done:
                    got[pending.pop(t)] = img
            assert steps < 20000
        for i, r in enumerate(ref):
            assert np.array_equal(got[i].view(np.uint32), r.view(np.uint32)), f"job {i}"
        assert dev.serve_info()["jobs_active"] == 0
    finally:
        dev.serve_configure(loop=True, sample_buffer_mb=4096)
