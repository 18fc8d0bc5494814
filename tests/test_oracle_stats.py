"""Pins of the oracle's per-path estimator (the one the HIP path reproduces
bit for bit) that do not come from a byte-identical reference output:

  * its IEEE-only math set (fm_sincos_turns, fm_exp, fm_cmf) against float64
    numpy and against the reference's CIE matching functions (glibc expf);
  * next-event estimation against a closed form for a rectangular
    Lambertian emitter (depth 1: camera hit + shadow rays only), with the
    reference CPU's light-cosine rule;
  * its image means against the branched estimator, which reproduces the
    reference CPU renderer byte for byte (code/Renderer.cpp:468-889): both
    estimate the same truncated path integral (one path per sample vs
    `samples` children per diffuse hit), so per-channel means must agree
    within 3 standard errors, the error estimated from independent seeds.
"""
import math
import os

import numpy as np
import pytest

import oracle_lib as ol
import scene_io

HERE = os.path.dirname(os.path.abspath(__file__))
SCENES = os.path.join(os.path.dirname(HERE), "pcr_amd", "scenes")


# ------------------------------------------------------------ fm math ------
def test_fm_sincos_turns_error_bound():
    """sin/cos(2*pi*t) for t in [0, 1] (NumGen draws, 1.0 included) within
    3e-7 absolute (about 2.5 ulp at 1) of float64."""
    L = ol.lib()
    t = np.concatenate([np.linspace(0.0, 1.0, 40001), np.array([1.0, 0.125, 0.375, 0.999999940395])]).astype(np.float32)
    out = np.zeros(2, np.float32)
    got = np.zeros((len(t), 2))
    for i, x in enumerate(t):
        L.orc_fm_sincos_turns(float(x), ol.fp(out))
        got[i] = out
    a = 2.0 * np.pi * t.astype(np.float64)
    assert np.max(np.abs(got[:, 0] - np.sin(a))) < 3e-7
    assert np.max(np.abs(got[:, 1] - np.cos(a))) < 3e-7
    # quadrant points are exact
    L.orc_fm_sincos_turns(0.25, ol.fp(out))
    assert out[0] == 1.0 and out[1] == 0.0
    L.orc_fm_sincos_turns(0.0, ol.fp(out))
    assert out[0] == 0.0 and out[1] == 1.0


def test_fm_exp_error_bound():
    """exp(x) within 2.5e-7 relative of float64 over the normal range, 0 and
    inf beyond it, and denormal results within one denormal step."""
    L = ol.lib()
    x = np.concatenate([np.linspace(-87.0, 88.0, 30001), np.linspace(-1.0, 1.0, 2001)]).astype(np.float32)
    got = np.array([L.orc_fm_exp(float(v)) for v in x], np.float64)
    want = np.exp(x.astype(np.float64))
    assert np.max(np.abs(got / want - 1.0)) < 2.5e-7
    assert L.orc_fm_exp(0.0) == 1.0
    assert L.orc_fm_exp(-104.0) == 0.0 and math.isinf(L.orc_fm_exp(89.0))
    xd = np.linspace(-102.0, -88.0, 201).astype(np.float32)
    gd = np.array([L.orc_fm_exp(float(v)) for v in xd], np.float64)
    wd = np.exp(xd.astype(np.float64))
    assert np.all(np.abs(gd - wd) <= np.maximum(2.5e-7 * wd, 2.0 ** -149))


def test_fm_cmf_matches_reference_cmf(golden):
    """The Wyman CMFs built on fm_exp against the reference's (glibc expf)
    values on 872 wavelengths: within 1e-6 absolute."""
    g = np.load(os.path.join(golden, "spectral.npz"))
    L = ol.lib()
    out = np.zeros(3, np.float32)
    got = np.zeros_like(g["cie_cmf"])
    for i, lam in enumerate(g["cie_lambda"]):
        L.orc_fm_cmf(float(lam), ol.fp(out))
        got[i] = out
    assert np.max(np.abs(got - g["cie_cmf"])) < 1e-6


# ------------------------------------------- closed-form direct lighting ---
# The reference CPU's light cosine is taken with the unnormalised light vector
# (code/Renderer.cpp:629,864: sampleN.dot(Li * -1), Li = sampleP - hit), so
# its next-event term integrates cos(theta) cos(theta') / d over the light
# instead of / d^2.  For a light parallel to the floor at height h,
# cos(theta) = cos(theta') = h/d, and the expected value of a depth-1 pixel is
#     (rho / pi) * Le * h * Omega(P),
# Omega(P) the solid angle of the light seen from the floor point P.  (The
# reference GPU backends normalise, code/Gpu/Opengl/OpenglRenderer.cpp:867;
# the CPU renderer is the oracle, so the CPU rule is the one reproduced.)
def corner_solid_angle(a, b, h):
    """Solid angle of an a-by-b rectangle seen from height h above one corner."""
    return math.atan(a * b / (h * math.sqrt(h * h + a * a + b * b)))


def rect_solid_angle(px, pz, x0, x1, z0, z1, h):
    """Solid angle of the rectangle [x0,x1] x [z0,z1] at height h seen from the
    floor point (px, 0, pz), by signed corner superposition."""
    def G(dx, dz):
        return math.copysign(1.0, dx) * math.copysign(1.0, dz) * corner_solid_angle(abs(dx), abs(dz), h)
    return G(x1 - px, z1 - pz) - G(x0 - px, z1 - pz) - G(x1 - px, z0 - pz) + G(x0 - px, z0 - pz)


def test_rect_solid_angle_against_quadrature():
    """The closed form itself, against midpoint quadrature of h/d^3 over the
    rectangle, for points under it and outside it."""
    x0, x1, z0, z1, h = -30.0, 30.0, -120.0, -40.0, 100.0
    n = 400
    xs = x0 + (np.arange(n) + 0.5) * (x1 - x0) / n
    zs = z0 + (np.arange(n) + 0.5) * (z1 - z0) / n
    X, Z = np.meshgrid(xs, zs)
    dA = (x1 - x0) * (z1 - z0) / (n * n)
    for px, pz in ((0.0, -80.0), (10.0, -50.0), (45.0, 20.0), (-80.0, -200.0)):
        d = np.sqrt((X - px) ** 2 + (Z - pz) ** 2 + h * h)
        quad = float(np.sum(h / d ** 3) * dA)
        assert abs(rect_solid_angle(px, pz, x0, x1, z0, z1, h) - quad) < 1e-5 * quad


def test_direct_lighting_closed_form():
    """Depth 1, RGB: a diffuse floor (albedo rho) under a 60 x 80 light of
    radiance Le at height 100, facing down.  Each pixel's expected value is
    (rho/pi) * Le * h * Omega(P) for its floor hit point P.  Checked per pixel
    (|z| < 4.5 over 192 pixels) and over the crop (|z| < 3), the standard
    error from 32 independent seeds; and the branched estimator (the
    reference CPU renderer) gives the same values."""
    L = ol.lib()
    rho, Le, h = 0.6, 7.0, 100.0
    x0, x1, z0, z1 = -30.0, 30.0, -120.0, -40.0
    sc = ol.OracleScene()
    L.orc_scene_add_material(sc.h, ol.fp(ol.f32([rho] * 3)), ol.fp(ol.f32([0, 0, 0])), 0, 0, 1.0, 0.0, None, None)
    L.orc_scene_add_material(sc.h, ol.fp(ol.f32([0, 0, 0])), ol.fp(ol.f32([Le] * 3)), 0, 0, 1.0, 0.0, None, None)
    # floor y = 0 (large parallelogram); light with u x v pointing to -y
    L.orc_scene_add_plane(sc.h, ol.fp(ol.f32([-5000, 0, 5000])), ol.fp(ol.f32([10000, 0, 0])),
                          ol.fp(ol.f32([0, 0, -10000])), 0, 0)
    L.orc_scene_add_plane(sc.h, ol.fp(ol.f32([x0, h, z0])), ol.fp(ol.f32([x1 - x0, 0, 0])),
                          ol.fp(ol.f32([0, 0, z1 - z0])), 1, 1)
    cam = (0.0, 40.0, 300.0)
    L.orc_scene_set_camera(sc.h, ol.fp(ol.f32(cam)), 40.0)
    assert L.orc_scene_finalize(sc.h) == 0
    W = H = 64
    crop = (24, 36, 16, 12)
    cx, cy, cw, ch = crop
    # floor hit point of each pixel's centre ray (the camera model of
    # Renderer.cpp:148-150,190-200), found with the oracle's own intersector
    scale = math.tan(math.pi / 180.0 * 40.0 * 0.5)
    rays = []
    for py in range(cy, cy + ch):
        for px in range(cx, cx + cw):
            x = (2.0 * (px + 0.5) / W - 1.0) * scale
            y = -(2.0 * (py + 0.5) / H - 1.0) * scale
            d = np.array([x, y, -1.0])
            rays.append(list(cam) + list(d / np.linalg.norm(d)))
    hits = sc.intersect(np.array(rays))
    assert np.all(hits["kind"] == 1) and np.all(hits["prim"] == 1)  # plane 1 = floor (lights first)
    om = np.array([rect_solid_angle(float(p[0]), float(p[2]), x0, x1, z0, z1, h) for p in hits["hit"]])
    want = (rho / math.pi * Le * h * om).reshape(ch, cw)
    K, spp = 32, 128
    est = np.stack([sc.render_paths(W, H, 1, spp, 4, ol.spread_seed(1000 + k), spectral=False, crop=crop)[0][..., 0] / spp
                    for k in range(K)])
    mean, se = est.mean(0), est.std(0, ddof=1) / math.sqrt(K)
    z = (mean - want) / se
    assert np.max(np.abs(z)) < 4.5, np.max(np.abs(z))
    tot = est.mean(axis=(1, 2))
    zt = (tot.mean() - want.mean()) / (tot.std(ddof=1) / math.sqrt(K))
    assert abs(zt) < 3.0, zt
    # the branched estimator: one path per pixel with 32 shadow rays, 64 seeds
    br = np.stack([sc.render_branched(W, H, 1, 1, 32, ol.spread_seed(3000 + k))[0][cy:cy + ch, cx:cx + cw, 0] for k in range(64)])
    zb = (br.mean() - want.mean()) / (br.mean(axis=(1, 2)).std(ddof=1) / 8.0)
    assert abs(zb) < 3.0, zb


# ---------------------------------------- per-path vs branched estimator ---
def scene(name):
    return scene_io.build_oracle_scene(scene_io.parse_scene(os.path.join(SCENES, name + ".json")))


# name -> (scene, depth, branched samples, per-path spp, flags)
MEAN_CASES = {
    "c0-spectral-d4": ("cornell-c0", 4, 4, 16, dict(spectral=True)),
    "c1-spectral-d6": ("cornell-c1", 6, 2, 16, dict(spectral=True)),
    "c1-rgb-mis-rr-strat-d5": ("cornell-c1", 5, 2, 16, dict(spectral=False, mis=True, russian=True,
                                                              stratified=True)),
    "c4-glass-spectral-d5": ("closed-c4", 5, 2, 16, dict(spectral=True)),
}


@pytest.mark.parametrize("case", sorted(MEAN_CASES))
def test_per_path_mean_equals_branched_mean(case):
    name, depth, samples, spp, flags = MEAN_CASES[case]
    sc = scene(name)
    W = H = 48
    K = 24
    spectral = flags.get("spectral", False)
    b = []
    p = []
    for k in range(K):
        lin, _ = sc.render_branched(W, H, depth, samples, 4, ol.spread_seed(5000 + k), **flags)
        b.append(lin.reshape(-1, 3).astype(np.float64).mean(0))
        acc, _ = sc.render_paths(W, H, depth, spp, 4, ol.spread_seed(7000 + k), **flags)
        m = acc.reshape(-1, 3).mean(0) / spp
        p.append(ol.xyz_to_srgb(m) if spectral else m)
    b, p = np.array(b), np.array(p)
    se = np.sqrt(b.var(0, ddof=1) / K + p.var(0, ddof=1) / K)
    z = (b.mean(0) - p.mean(0)) / se
    assert np.all(np.abs(z) < 3.0), (b.mean(0), p.mean(0), z)
