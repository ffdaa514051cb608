"""Checks against the committed golden fixtures, which hold the REAL
reference's outputs (written by tests/golden/make_golden.py): the oracle on
the CPU, and the device path on the GPU."""
import os

import numpy as np
import pytest

import xtracer_amd as xa
from oracle.oracle import OracleScene
from xtracer_amd import scenes
from tests.helpers import edge_band, plane_point, soup_to_object_local

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
SCENES = {"spheres": ("spheres", {}), "terrain": ("terrain", {"n": 41}), "ico3": ("icosphere", {"subdiv": 3})}


def _scene(name):
    kind, kw = SCENES[name]
    return xa.Scene(scenes.SCENES[kind][0](**kw)), scenes.camera_dict(**scenes.SCENES[kind][1])


def _load(prefix, name):
    return np.load(os.path.join(HERE, f"{prefix}_{name}.npz"))


@pytest.mark.parametrize("name", list(SCENES))
def test_oracle_trace_equals_reference_fixture(name):
    g = _load("ref_trace", name)
    s, _ = _scene(name)
    t, obj, _, pt, _ = OracleScene(s).trace(g["o"].astype(np.float64), g["d"].astype(np.float64))
    np.testing.assert_array_equal(obj, g["obj"])
    np.testing.assert_array_equal(t, g["t"])
    np.testing.assert_array_equal(pt[obj >= 0], g["point"][obj >= 0])


@pytest.mark.parametrize("name", ["spheres", "terrain"])
def test_oracle_render_equals_reference_fixture(name):
    g = _load("ref_render", name)
    s, cam = _scene(name)
    orc = OracleScene(s)
    for integ, key in ((0, "stencil"), (1, "depth")):
        img, _ = orc.render(integ, int(g["width"]), int(g["height"]), cam, samples=int(g["spp"]))
        assert np.array_equal(img.view(np.uint32), g[key].view(np.uint32)), key


@pytest.mark.gpu
@pytest.mark.parametrize("accel", [xa.ACCEL_KDTREE, xa.ACCEL_OCTREE])
@pytest.mark.parametrize("name", list(SCENES))
def test_device_trace_vs_reference_fixture(device, name, accel):
    g = _load("ref_trace", name)
    s, _ = _scene(name)
    sid = device.upload(s)
    t, prim, u, v, flags = device.trace(sid, g["o"], g["d"], accel)
    arr = s.arrays()
    obj, loc = soup_to_object_local(arr, prim)
    # the oracle equals these fixtures bit for bit (obj, t, point); it names the
    # reference's primitive inside the object
    _, oobj, oprim, opt, _ = OracleScene(s).trace(g["o"].astype(np.float64), g["d"].astype(np.float64))
    np.testing.assert_array_equal(oobj, g["obj"])
    same = (obj == g["obj"]) & ((obj < 0) | (loc == oprim))
    mism = np.flatnonzero(~same)
    if len(mism):
        # every disagreement lies in the edge band of the device's or the
        # reference's triangle
        band = edge_band(arr, prim[mism], plane_point(arr, prim[mism], g["o"][mism], g["d"][mism]))
        inv = {(int(a), int(b)): k for k, (a, b) in enumerate(zip(arr["prims"]["object"], arr["prim_local"]))}
        rp = np.array([inv.get((int(a), int(b)), 0xFFFFFFFF) for a, b in zip(oobj[mism], oprim[mism])], np.uint64)
        band |= edge_band(arr, rp, g["point"][mism])
        assert band.all(), f"disagreements outside the edge band: rays {mism[~band][:5]}"
    assert same.mean() >= 0.999, f"{(~same).sum()} of {len(same)} disagree"
    hit = same & (g["obj"] >= 0)
    rel = np.abs(t[hit] - g["t"][hit]) / g["t"][hit]
    assert rel.max() <= 1e-5


@pytest.mark.gpu
@pytest.mark.parametrize("accel", [xa.ACCEL_KDTREE, xa.ACCEL_OCTREE])
@pytest.mark.parametrize("name", ["spheres", "terrain"])
def test_device_render_equals_reference_fixture(device, name, accel):
    g = _load("ref_render", name)
    s, _ = _scene(name)
    sid = device.upload(s)
    for integ, key in ((xa.STENCIL, "stencil"), (xa.DEPTH, "depth")):
        job = xa.make_job(sid, integ, int(g["width"]), int(g["height"]), s.camera(), samples=int(g["spp"]),
                          accel=accel)
        img, _ = device.render(job)
        bad = (img.view(np.uint32) != g[key].view(np.uint32)).any(axis=2)
        assert not bad.any(), f"{key}: {bad.sum()} pixels differ"


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["spheres", "terrain"])
def test_device_monte_carlo_vs_reference_fixture(device, name):
    """8 device renders at 64 spp against the reference's 2048-spp image:
    per-pixel agreement within the 99% bound built from the device's spread."""
    g = _load("ref_mc", name)
    s, _ = _scene(name)
    sid = device.upload(s)
    imgs = []
    for k in range(8):
        job = xa.make_job(sid, int(g["integrator"]), int(g["width"]), int(g["height"]), s.camera(), samples=64,
                          rdepth=int(g["rdepth"]), seed=500 + k)
        imgs.append(device.render(job)[0][..., :3])
    imgs = np.stack(imgs)
    mean = imgs.mean(0)
    sd_img = imgs.std(0, ddof=1)
    # the spread comes from 8 images: 99% two-sided quantile of Student's t, 7 dof
    bound = _mc_bound(mean, sd_img ** 2 * 64, 8 * 64, int(g["spp"]), z=3.499)
    inside = np.abs(mean - g["image"][..., :3]) <= bound
    assert inside.mean() >= 0.98, f"{inside.mean():.3f} inside the 99% bound"


# The 4096-spp reference images (ref_mc4096_*.npz): ao on the spheres,
# pathtracer on the terrain and on the displaced icosphere with the thin-lens
# camera, Phong floor and area lights at rdepth 8.
MC4096 = ["spheres", "terrain", "ico3"]


def _sample_var(var, mean, n):
    """Per-sample variance, where the n samples of a channel were all equal
    that of an event of probability 3 / n (rule of three: not seen in n
    samples with 95% confidence) changing the value by its size."""
    return np.where(var > 0, var, np.maximum(mean ** 2, 1e-6) * 3.0 / n)


def _mc_bound(mean, var, n, n_ref, z=2.576):
    """Per-channel 99% bound of (estimate over n samples) - (reference over
    n_ref samples), both with per-sample variance var."""
    v = _sample_var(var, mean, n)
    return z * np.sqrt(v / n + v / n_ref) + 1e-4


@pytest.mark.parametrize("name", MC4096)
def test_oracle_monte_carlo_vs_4096spp_reference(name):
    """The oracle at 512 spp (its own per-sample variance) against the
    reference's 4096-spp image: at least 97% of channels inside the 99%
    bound, and no channel beyond 3 times it."""
    g = _load("ref_mc4096", name)
    s, cam = _scene(name)
    w, h, n = int(g["width"]), int(g["height"]), 512
    _, _, mean, m2 = OracleScene(s).render(int(g["integrator"]), w, h, cam, samples=n, rdepth=int(g["rdepth"]),
                                          seed=77, stats=True)
    var = m2 / (n - 1)
    bound = _mc_bound(mean, var, n, int(g["spp"]))
    dev = np.abs(mean - g["image"][..., :3])
    assert (dev <= bound).mean() >= 0.97, f"{(dev <= bound).mean():.4f} inside the 99% bound"
    assert (dev <= 3 * bound).all()


@pytest.mark.gpu
@pytest.mark.parametrize("name", MC4096)
def test_device_monte_carlo_vs_4096spp_reference(device, name):
    """8 device renders of 128 spp against the reference's 4096-spp image: at
    least 98% of channels inside the 99% bound (Student's t on the 8 images'
    spread, 7 degrees of freedom, plus the reference's own variance scaled
    from it)."""
    g = _load("ref_mc4096", name)
    s, _ = _scene(name)
    sid = device.upload(s)
    spp = 128
    imgs = np.stack([device.render(xa.make_job(sid, int(g["integrator"]), int(g["width"]), int(g["height"]), s.camera(),
                                               samples=spp, rdepth=int(g["rdepth"]), seed=900 + k))[0][..., :3]
                     for k in range(8)])
    mean, sd = imgs.mean(0), imgs.std(0, ddof=1)
    # per-sample variance from the 8 images' spread; Student's t, 7 dof
    bound = _mc_bound(mean, sd ** 2 * spp, 8 * spp, int(g["spp"]), z=3.499)
    inside = np.abs(mean - g["image"][..., :3]) <= bound
    assert inside.mean() >= 0.98, f"{inside.mean():.4f} inside the 99% bound"
    device.release(sid)
