"""Oracle optics vs the compiled reference and tests/unit/optics.cpp."""
import numpy as np

import oracle_lib as ol


def bits(a):
    return np.asarray(a, np.float32).view(np.uint32)


def test_dielectric_bitexact(golden):
    g = np.load(golden + "/optics.npz")
    L = ol.lib()
    out = np.zeros(6, np.float32)
    for i, r in enumerate(g["inputs"]):
        L.orc_dielectric(ol.fp(r[0:3].copy()), ol.fp(r[3:6].copy()), ol.fp(r[6:9].copy()), int(r[9] != 0),
                         float(r[10]), float(r[11]), ol.fp(out))
        assert np.array_equal(bits(out[:3]), bits(g["dir"][i])), i
        assert np.array_equal(bits(out[3:]), bits(g["origin"][i])), i
        cos_i = np.float32(-(np.float32(r[0] * r[3]) + np.float32(r[1] * r[4]) + np.float32(r[2] * r[5])))
        s = L.orc_schlick(float(abs(np.float32(-np.dot(r[0:3], r[3:6])))), 1.0, float(r[10]))
        assert abs(s - g["schlick"][i]) <= 1e-6 * max(1, abs(g["schlick"][i]))
        c = L.orc_cauchy(float(r[10]), 0.42, float(np.float32(400 + 300 * r[11])))
        assert bits(c) == bits(g["cauchy"][i])


def test_unit_values():
    L = ol.lib()
    assert abs(L.orc_schlick(1.0, 1.0, 1.5) - 0.04) < 1e-4
    assert abs(L.orc_schlick(1.0, 1.5, 1.5)) < 1e-6
    assert abs(L.orc_schlick(0.0, 1.0, 1.5) - 1) < 1e-3
    n400, n700 = L.orc_cauchy(1.5, 0.013, 400.0), L.orc_cauchy(1.5, 0.013, 700.0)
    assert 3e-4 < n400 - n700 < 1e-3
    out = np.zeros(6, np.float32)
    d = ol.f32([np.sqrt(0.75), -0.5, 0])
    L.orc_dielectric(ol.fp(d), ol.fp(ol.f32([0, 1, 0])), ol.fp(ol.f32([0, 0, 0])), 0, 1.5, 0.99, ol.fp(out))
    assert out[1] > 0 and out[4] > 0  # TIR reflects and offsets along +N
    L.orc_dielectric(ol.fp(ol.f32([0.6, -0.8, 0])), ol.fp(ol.f32([0, 1, 0])), ol.fp(ol.f32([0, 0, 0])), 1, 1.5,
                     0.0, ol.fp(out))
    assert abs(out[0] - 0.6) < 1e-4 and abs(out[1] - 0.8) < 1e-4
