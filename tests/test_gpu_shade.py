"""IMaterial::shade on the device (xt_shade) against the oracle, whose
restatement equals the reference materials' own shade() bit for bit
(tests/test_shade.py). Lambert, emissive and dielectric are bit-exact; Phong
and Blinn-Phong raise max(0, cos) to the exponent in FP64 where the reference
uses long double before rounding to float, so they may differ in the last
float place."""
import numpy as np
import pytest

import xtracer_amd as xa
from oracle import oracle as O
from tests.test_shade import OBJECTS, random_queries, shade_scene


@pytest.mark.gpu
def test_device_shade_equals_oracle(device, tmp_path):
    s = xa.Scene(shade_scene(tmp_path))
    sid = device.upload(s)
    q, em, io = random_queries(20000, len(s.objects()), 5)
    a, ra = device.shade(sid, q, em, io)
    b, rb = O.OracleScene(s).shade(q, em, io)
    np.testing.assert_array_equal(ra, rb)
    names = [o["name"] for o in s.objects()]
    kind = np.array([OBJECTS[int(names[int(k)][1:])] if int(names[int(k)][1:]) < len(OBJECTS) else "none"
                     for k in q[:, 15]])
    ulps = np.abs(a.view(np.int32).astype(np.int64) - b.view(np.int32).astype(np.int64)).max(axis=1)
    exact = ~np.isin(kind, ["pho", "bp"])
    assert (ulps[exact] == 0).all(), f"{(ulps[exact] != 0).sum()} exact-material queries differ"
    assert ulps[~exact].max() <= 1
    assert (ulps[~exact] == 0).mean() >= 0.99
    device.release(sid)
