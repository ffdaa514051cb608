"""Closest-hit parity of the device traversal (FP32 + FP64 re-test) against the
oracle's FP64 restatement of Scene::intersection, on identical FP32 rays."""
import numpy as np
import pytest

import xtracer_amd as xa
from oracle.oracle import OracleScene
from tests.helpers import camera_rays, mixed_rays, scene_path, soup_to_object_local

SCENES = [("spheres", {}), ("icosphere", {"subdiv": 4}), ("strands", {"strands": 400}), ("terrain", {"n": 41})]


def _bary(p, a, b, c):
    v0, v1, v2 = b - a, c - a, p - a
    d00 = (v0 * v0).sum(1)
    d01 = (v0 * v1).sum(1)
    d11 = (v1 * v1).sum(1)
    d20 = (v2 * v0).sum(1)
    d21 = (v2 * v1).sum(1)
    den = d00 * d11 - d01 * d01
    v = (d11 * d20 - d01 * d21) / den
    w = (d00 * d21 - d01 * d20) / den
    return np.stack([1 - v - w, v, w], 1)


def _edge_band(arr, prim, point, tol=1e-3):
    """True where `point` lies within `tol` (barycentric) of an edge of soup prim `prim`."""
    out = np.zeros(len(prim), bool)
    p64 = arr["prims64"]
    ok = prim < len(p64)
    if ok.any():
        q = p64[prim[ok].astype(np.int64)]
        tri = (arr["prims"]["tf"][prim[ok].astype(np.int64)] & 15) == 0
        bc = _bary(point[ok], q[:, 0:3], q[:, 3:6], q[:, 6:9])
        out[np.flatnonzero(ok)] = tri & (np.abs(bc).min(1) < tol)
    return out


@pytest.mark.gpu
@pytest.mark.parametrize("accel", [xa.ACCEL_KDTREE, xa.ACCEL_OCTREE])
@pytest.mark.parametrize("kind,kw", SCENES)
def test_hit_parity(device, kind, kw, accel):
    s = xa.Scene(scene_path(kind, **kw))
    sid = device.upload(s)
    arr = s.arrays()
    orc = OracleScene(s)
    o1, d1 = camera_rays(s.camera(), 160, 120)
    o2, d2 = mixed_rays(None, 1 << 15, seed=7)
    o = np.concatenate([o1.astype(np.float32), o2])
    d = np.concatenate([d1.astype(np.float32), d2])
    t, prim, u, v, flags = device.trace(sid, o, d, accel)
    tr, obr, pr, ptr, nr = orc.trace(o.astype(np.float64), d.astype(np.float64))
    obj, loc = soup_to_object_local(arr, prim)
    same = (obj == obr) & (loc == pr)
    n = len(o)
    mism = np.flatnonzero(~same)
    # every disagreement must sit in the edge band of one of the two primitives
    if len(mism):
        gpu_pt = o[mism].astype(np.float64) + d[mism].astype(np.float64) * np.where(np.isfinite(t[mism]), t[mism], 0)[:, None]
        band = _edge_band(arr, prim[mism], gpu_pt)
        # the oracle's primitive as a soup index
        first = {k: ob["first_prim"] for k, ob in enumerate(s.objects())}
        inv = {}
        for sidx, (ob_, lc) in enumerate(zip(arr["prims"]["object"], arr["prim_local"])):
            inv[(int(ob_), int(lc))] = sidx
        oprim = np.array([inv.get((int(a), int(b)), 0xFFFFFFFF) for a, b in zip(obr[mism], pr[mism])], np.uint64)
        band |= _edge_band(arr, oprim, ptr[mism])
        assert band.all(), f"{(~band).sum()} disagreements outside the edge band, e.g. ray {mism[~band][:5]}"
    assert same.mean() >= 0.9999, f"{kind}: hit-id agreement {same.mean():.6f} ({len(mism)} of {n})"
    hit = same & (obr >= 0)
    rel = np.abs(t[hit].astype(np.float64) - tr[hit]) / np.maximum(tr[hit], 1e-30)
    assert rel.max() <= 1e-5, f"t relative error {rel.max():.3g}"
    assert np.isinf(t[same & (obr < 0)]).all()


@pytest.mark.gpu
def test_kdtree_octree_agree(device):
    s = xa.Scene(scene_path("icosphere", subdiv=5))
    sid = device.upload(s)
    o, d = mixed_rays(None, 1 << 16, seed=3)
    a = device.trace(sid, o, d, xa.ACCEL_KDTREE)
    b = device.trace(sid, o, d, xa.ACCEL_OCTREE)
    assert np.array_equal(a[1], b[1])
    # t may differ by FP32 rounding where one traversal sent the ray to the
    # FP64 re-test and the other did not
    np.testing.assert_allclose(a[0], b[0], rtol=1e-6)
