"""Watertightness at the benchmark's size: a ray from outside a closed mesh
enters it through a front face, a ray from inside leaves through a back face.
A traversal or triangle test that let rays slip between adjacent triangles
would show up as first hits on the wrong side (the reference's FP64
area-sum test has no such gaps). The displaced icosphere is star-shaped
about its centre with outward winding."""
import numpy as np
import pytest

import xtracer_amd as xa
from tests.helpers import camera_rays, scene_path


def _ico(s):
    for ob in s.objects():
        if ob["name"] == "ico":
            return ob
    raise AssertionError("no ico object")


def _facing(arr, prim, d):
    p = arr["prims64"][prim.astype(np.int64)]
    n = np.cross(p[:, 3:6] - p[:, 0:3], p[:, 6:9] - p[:, 0:3])
    n /= np.linalg.norm(n, axis=1)[:, None]
    return np.einsum("ij,ij->i", n, d.astype(np.float64))


@pytest.mark.gpu
@pytest.mark.parametrize("subdiv", [6, 8])
def test_closed_mesh_entry_and_exit(device, subdiv):
    s = xa.Scene(scene_path("icosphere", subdiv=subdiv))
    sid = device.upload(s)
    arr = s.arrays()
    ob = _ico(s)
    lo, hi = ob["first_prim"], ob["first_prim"] + ob["prim_count"]
    verts = arr["prims64"][lo:hi, 0:3]
    centre = 0.5 * (verts.min(0) + verts.max(0))
    rng = np.random.default_rng(subdiv)
    # outside: the scene camera's primary rays and rays aimed at the mesh from a shell around it
    o1, d1 = camera_rays(s.camera(), 1920, 1080)
    n2 = 1 << 20
    u = rng.normal(size=(n2, 3))
    u /= np.linalg.norm(u, axis=1)[:, None]
    o2 = centre + 3.0 * u
    aim = centre + 0.9 * rng.uniform(-1, 1, (n2, 3))
    d2 = aim - o2
    d2 /= np.linalg.norm(d2, axis=1)[:, None]
    # inside: rays from near the centre in every direction
    o3 = centre + 0.05 * rng.uniform(-1, 1, (n2, 3))
    d3 = rng.normal(size=(n2, 3))
    d3 /= np.linalg.norm(d3, axis=1)[:, None]
    for o, d, sign, what in ((o1, d1, -1.0, "camera"), (o2, d2, -1.0, "outside"), (o3, d3, 1.0, "inside")):
        o32, d32 = o.astype(np.float32), d.astype(np.float32)
        t, prim, _, _, _ = device.trace(sid, o32, d32)
        on = (prim >= lo) & (prim < hi)
        if what == "inside":
            assert on.all(), f"{(~on).sum()} rays from inside the mesh escaped it"
        cos = _facing(arr, prim[on], d32[on])
        wrong = sign * cos < -1e-4  # clearly on the wrong side (not grazing)
        assert on.sum() > 0.2 * len(o)
        assert not wrong.any(), (f"{what}: {wrong.sum()} of {on.sum()} first hits on the wrong face, "
                                 f"e.g. rays {np.flatnonzero(on)[wrong][:5]}")
    device.release(sid)
