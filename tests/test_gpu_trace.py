"""Closest-hit parity of the device traversal (FP32 + FP64 re-test) against the
oracle's FP64 restatement of Scene::intersection, on identical FP32 rays."""
import numpy as np
import pytest

import xtracer_amd as xa
from oracle.oracle import OracleScene
from tests.helpers import (bary as _bary, camera_rays, edge_band as _edge_band, mixed_rays, plane_point as _plane_point,
                           record_normals, scene_path, soup_to_object_local)

SCENES = [("spheres", {}), ("icosphere", {"subdiv": 4}), ("strands", {"strands": 400}), ("terrain", {"n": 41})]


@pytest.mark.gpu
@pytest.mark.parametrize("accel,kd_build", [(xa.ACCEL_KDTREE, "host"), (xa.ACCEL_OCTREE, "host"),
                                            (xa.ACCEL_KDTREE, "device")])
@pytest.mark.parametrize("kind,kw", SCENES)
def test_hit_parity(device, kind, kw, accel, kd_build):
    """kd_build "device": the kd-tree rebuilt on the device."""
    s = xa.Scene(scene_path(kind, **kw))
    sid = device.upload(s)
    if kd_build == "device":
        device.build_kd(sid)
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
        band = _edge_band(arr, prim[mism], _plane_point(arr, prim[mism], o[mism], d[mism]))
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
    # barycentrics of agreed triangle hits against the oracle's FP64 hit point
    soup = hit & (prim < len(arr["prims64"]))
    tri = np.zeros(n, bool)
    tri[soup] = (arr["prims"]["tf"][prim[soup].astype(np.int64)] & 15) == 0
    if tri.any():
        q = arr["prims64"][prim[tri].astype(np.int64)]
        bc = _bary(ptr[tri], q[:, 0:3], q[:, 3:6], q[:, 6:9])
        err = np.maximum(np.abs(u[tri] - bc[:, 1]), np.abs(v[tri] - bc[:, 2]))
        assert err.max() <= 1e-5, f"barycentric error {err.max():.3g}"
    # normals that the hit records give, within 1e-5 relative
    nd = record_normals(s, arr, o, d, t, prim, u, v)
    dn = np.linalg.norm(nd[hit] - nr[hit], axis=1) / np.maximum(np.linalg.norm(nr[hit], axis=1), 1e-30)
    assert dn.max() <= 1e-5, f"normal relative error {dn.max():.3g}"


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
