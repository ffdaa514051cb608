"""Texture, gradient, ERP and cubemap samplers on the device against the
oracle, which matches the reference bit for bit on the same scenes.

A sampler reads texels at float coordinates; the planes', triangles' and
meshes' texcoords and the cubemap mapping use only the basic operations and
come out bit-exact. A sphere's texcoord takes asin, and the ERP sampler
acos / atan2, which the device's math library may round differently in the
last bit; where such a coordinate lands on a texel edge the two pick
neighbouring texels, so a small share of pixels may differ by a texel."""
import numpy as np
import pytest

import xtracer_amd as xa
from oracle.oracle import OracleScene
from tests.helpers import camera_dict_of, same_stream_check, texture_scene


def _render(device, path, cam, integrator, accel=xa.ACCEL_KDTREE, W=96, H=72, spp=4, stats=False, **kw):
    s = xa.Scene(path)
    sid = device.upload(s)
    try:
        img, st = device.render(xa.make_job(sid, integrator, W, H, s.camera(), samples=spp, seed=5, accel=accel,
                                            **kw))
    finally:
        device.release(sid)
    out = OracleScene(s).render(integrator, W, H, camera_dict_of(cam), samples=spp, seed=5,
                                rdepth=kw.get("rdepth", 100), stats=stats)
    return (img, st) + tuple(out)


@pytest.mark.gpu
@pytest.mark.parametrize("cam", ["pinhole", "erp", "cubemap"])
@pytest.mark.parametrize("accel", [xa.ACCEL_KDTREE, xa.ACCEL_OCTREE])
def test_material_samplers(device, cam, accel, tmp_path):
    """Emission integrator: emissive textures (nearest with flip_x and
    texcoords outside [0, 1], bilinear with and without flip_y, RGBA), a
    gradient and an ERP sampler read at each hit's texcoord."""
    path = texture_scene(str(tmp_path / "t.scn"), cam, "emit", "erp")
    img, st, ref, nq = _render(device, path, cam, xa.EMISSION, accel)
    assert st.queries == nq
    exact = (img.view(np.uint32) == ref.view(np.uint32)).all(axis=2)
    assert exact.mean() >= 0.99, f"only {exact.mean():.4f} of pixels bit-exact"
    assert (ref[..., :3] > 0).any(axis=2).sum() > 2000


@pytest.mark.gpu
@pytest.mark.parametrize("cam", ["pinhole", "erp", "cubemap"])
@pytest.mark.parametrize("env", ["erp", "cubemap", "gradient"])
def test_environment_samplers(device, cam, env, tmp_path):
    """Pathtracer rays that miss return the environment's sample, looked up
    along the path's FP32 ray direction (the oracle, like the reference,
    uses the FP64 one). The gradient is continuous in the direction: every
    pixel within 1e-6 relative. The cubemap sampler picks texels with the
    basic operations: with the pinhole and cubemap cameras its image is
    bit-exact. ERP lookups (acos / atan2) may land on a neighbouring texel
    for a few pixels."""
    path = texture_scene(str(tmp_path / "t.scn"), cam, "env", env)
    img, st, ref, nq = _render(device, path, cam, xa.PATHTRACER, W=64, H=48, rdepth=6, accum=xa.ACCUM_REFERENCE)
    assert st.queries == nq
    exact = (img.view(np.uint32) == ref.view(np.uint32)).all(axis=2)
    if env == "gradient":
        close = np.isclose(img, ref, rtol=1e-6, atol=0).all(axis=2)
        assert close.all(), f"{(~close).sum()} pixels differ by more than 1e-6"
    elif env == "cubemap" and cam != "erp":
        assert exact.all(), f"{(~exact).sum()} pixels differ"
    else:
        assert exact.mean() >= 0.99, f"only {exact.mean():.4f} of pixels bit-exact"
    assert len(np.unique(ref[..., :3].reshape(-1, 3), axis=0)) > 10


@pytest.mark.gpu
@pytest.mark.parametrize("cam,env", [("pinhole", "erp"), ("erp", "cubemap")])
def test_textured_pathtracer_same_stream(device, cam, env, tmp_path):
    """Textured lambert / phong / blinn-phong channels, a cubemap-sampled
    diffuse, a textured light and a sampled environment: with the same seed
    the device and the oracle agree on all but the FP32-decided pixels."""
    path = texture_scene(str(tmp_path / "t.scn"), cam, "lit", env)
    img, st, ref, nq, mean, m2 = _render(device, path, cam, xa.PATHTRACER, W=64, H=48, spp=8, rdepth=6, stats=True,
                                         accum=xa.ACCUM_REFERENCE)
    same_stream_check(img, ref, mean, m2, 8, 0.97)
    assert abs(st.queries - nq) <= 0.01 * nq


@pytest.mark.gpu
@pytest.mark.parametrize("accel", [xa.ACCEL_KDTREE, xa.ACCEL_OCTREE])
def test_extruded_mesh(device, accel, tmp_path):
    """A mesh pushed out by the extrude modifier's cubemap: stencil, depth
    and normal images bit-exact against the oracle."""
    from tests.helpers import extrude_scene
    from xtracer_amd.scenes import camera_dict
    path = extrude_scene(str(tmp_path / "x.scn"))
    s = xa.Scene(path)
    sid = device.upload(s)
    cam = camera_dict(position=(0.5, 2.0, -9.0), target=(0.0, 0.3, 0.0), fov=50.0)
    try:
        for integrator in (xa.STENCIL, xa.DEPTH, xa.NORMAL):
            img, st = device.render(xa.make_job(sid, integrator, 96, 72, s.camera(), samples=4, seed=5, accel=accel))
            ref, nq = OracleScene(s).render(integrator, 96, 72, cam, samples=4, seed=5)
            assert st.queries == nq
            assert np.array_equal(img.view(np.uint32), ref.view(np.uint32)), integrator
    finally:
        device.release(sid)


@pytest.mark.gpu
@pytest.mark.parametrize("accel", [xa.ACCEL_KDTREE, xa.ACCEL_OCTREE])
def test_external_objects(device, accel, tmp_path):
    """An OBJ object with MTL materials (an emissive shape textured through
    map_Ke, mirrored left to right; colour, emissive and Blinn-Phong shapes)
    beside gen(icosahedron) and gen(plane) meshes: stencil, normal and
    emission images bit-exact against the oracle (triangles only read
    textures here), the pathtracer on the same random stream."""
    from tests.helpers import external_scene
    path = external_scene(str(tmp_path / "e.scn"))
    for integrator in (xa.STENCIL, xa.NORMAL, xa.EMISSION):
        img, st, ref, nq = _render(device, path, "pinhole", integrator, accel)
        assert st.queries == nq
        exact = (img.view(np.uint32) == ref.view(np.uint32)).all(axis=2)
        assert exact.all(), f"integrator {integrator}: {(~exact).sum()} pixels differ"
        assert (ref[..., :3] > 0).any(axis=2).sum() > 100
    img, st, ref, nq, mean, m2 = _render(device, path, "pinhole", xa.PATHTRACER, accel, W=64, H=48, spp=8, rdepth=6,
                                         stats=True, accum=xa.ACCUM_REFERENCE)
    same_stream_check(img, ref, mean, m2, 8, 0.97)
