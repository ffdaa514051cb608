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
    se = np.sqrt(sd_img ** 2 / 8 + sd_img ** 2 * 64 / int(g["spp"]))
    # the spread comes from 8 images: 99% two-sided quantile of Student's t, 7 dof
    inside = np.abs(mean - g["image"][..., :3]) <= 3.499 * se + 1e-4
    assert inside.mean() >= 0.95, f"{inside.mean():.3f} inside the 99% bound"
