"""Writes the benchmark scenes (reference scene schema 1.0) and their SPDs.

Scenes (all in a 555-unit Cornell box, x right, y up, the open side facing the
camera at +z; the reference camera is fixed looking down -z):
  cornell-c0       quad walls as triangle pairs, two rotated diffuse boxes,
                   130x105 ceiling light with a flat emission SPD, wall
                   reflectances as smooth spectral curves.
  cornell-c1       the same box with a mirror sphere and a Cauchy-dispersive
                   glass sphere (IOR 1.5, B = 0.0042 um^2 -> cauchyB 0.42 in the
                   reference's units, ior(l) = ior + cauchyB*1e4/l^2).
  cornell-c2       the same box with a displaced icosphere (subdivision 6,
                   3-octave value noise, amplitude 0.15 r, 200 units across);
                   uses the "icosphere" primitive of this project's loader.
  closed-c3        closed six-wall box, 30x30 ceiling light, one diffuse sphere.
  closed-c4        closed six-wall box, 30x30 ceiling light, one glass sphere.

Coordinates are float32 values printed with 9 significant digits, so the JSON
fully determines every vertex.  Run: python pcr_amd/scenes/make_scenes.py
"""
import json
import math
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
L = 555.0


def f32(x):
    return float(np.float32(x))


def v(*xs):
    return [float("%.9g" % f32(x)) for x in xs]


def quad_tris(a, b, c, d, mat, light=False):
    """Two triangles (a,b,c) and (a,c,d)."""
    out = []
    for t in ((a, b, c), (a, c, d)):
        p = {"type": "triangle", "v0": v(*t[0]), "v1": v(*t[1]), "v2": v(*t[2]), "material": mat}
        if light:
            p["light"] = True
        out.append(p)
    return out


def walls(closed=False):
    P = []
    # floor y=0, ceiling y=L, back z=-L, left x=0 (red), right x=L (green)
    P += quad_tris((0, 0, 0), (L, 0, 0), (L, 0, -L), (0, 0, -L), "white")
    P += quad_tris((0, L, 0), (0, L, -L), (L, L, -L), (L, L, 0), "white")
    P += quad_tris((0, 0, -L), (L, 0, -L), (L, L, -L), (0, L, -L), "white")
    P += quad_tris((0, 0, 0), (0, 0, -L), (0, L, -L), (0, L, 0), "red")
    P += quad_tris((L, 0, 0), (L, L, 0), (L, L, -L), (L, 0, -L), "green")
    if closed:
        P += quad_tris((0, 0, 0), (0, L, 0), (L, L, 0), (L, 0, 0), "white")
    return P


def ceiling_light(w, d, y=L - 1.0):
    """Light quad centred under the ceiling, normal pointing down (-y)."""
    x0, x1 = (L - w) / 2, (L + w) / 2
    z0, z1 = -(L + d) / 2, -(L - d) / 2
    return quad_tris((x0, y, z0), (x1, y, z0), (x1, y, z1), (x0, y, z1), "light", light=True)


def box(cx, cz, sx, sy, sz, deg, mat):
    """Axis box rotated about y, resting on the floor; bottom face omitted."""
    a = math.radians(deg)
    ca, sa = math.cos(a), math.sin(a)

    def p(x, y, z):
        x -= sx / 2
        z -= sz / 2
        return (cx + ca * x + sa * z, y, cz - sa * x + ca * z)

    c = [p(x, y, z) for y in (0, sy) for z in (0, sz) for x in (0, sx)]
    # c index: (y, z, x) bits -> 4*y + 2*z + x
    P = []
    P += quad_tris(c[4], c[5], c[7], c[6], mat)   # top
    P += quad_tris(c[0], c[1], c[5], c[4], mat)   # z = 0 side
    P += quad_tris(c[2], c[6], c[7], c[3], mat)   # z = sz side
    P += quad_tris(c[0], c[4], c[6], c[2], mat)   # x = 0 side
    P += quad_tris(c[1], c[3], c[7], c[5], mat)   # x = sx side
    return P


def spd_curves():
    lam = np.arange(400, 701, 10, dtype=np.float64)
    white = 0.70 + 0.05 * (1.0 / (1.0 + np.exp(-(lam - 450.0) / 25.0)))
    red = 0.05 + 0.60 * (1.0 / (1.0 + np.exp(-(lam - 590.0) / 15.0)))
    green = 0.08 + 0.40 * np.exp(-0.5 * ((lam - 530.0) / 35.0) ** 2)
    flat = np.ones_like(lam)
    return lam, {"white": white, "red": red, "green": green, "flat": flat}


def write_spds():
    lam, curves = spd_curves()
    d = os.path.join(HERE, "spd", "pcr")
    os.makedirs(d, exist_ok=True)
    for name, vals in curves.items():
        with open(os.path.join(d, name + ".spd"), "w") as f:
            f.write("# %s: smooth synthetic curve for the benchmark scenes\n" % name)
            f.write("# wavelength_nm value\n")
            for l_, v_ in zip(lam, vals):
                f.write("%d %.4f\n" % (l_, v_))


def materials(extra=None, light_scale=15.0):
    m = {
        "white": {"albedo": [0.73, 0.73, 0.72], "albedo_spd": "pcr/white"},
        "red": {"albedo": [0.63, 0.06, 0.05], "albedo_spd": "pcr/red"},
        "green": {"albedo": [0.13, 0.45, 0.10], "albedo_spd": "pcr/green"},
        "light": {"emissive": [light_scale, light_scale, light_scale], "emissive_spd": "pcr/flat",
                  "emissive_scale": light_scale},
    }
    if extra:
        m.update(extra)
    return m


def scene(name, prims, mats, cam=(L / 2, L / 2, 800.0), fov=39.3, comment=""):
    return {"schema": "1.0", "name": name, "version": "1.0.0", "comment": comment,
            "camera": {"position": v(*cam), "fov": fov}, "materials": mats, "primitives": prims}


def main():
    write_spds()
    glass = {"albedo": [1.0, 1.0, 1.0], "transparent": True, "ior": 1.5, "cauchyB": 0.42}
    mirror = {"albedo": [1.0, 1.0, 1.0], "metallic": True}
    scenes = {}
    c0 = walls() + box(370.0, -170.0, 165, 165, 165, 18.0, "white") + \
        box(185.0, -370.0, 165, 330, 165, -15.0, "white") + ceiling_light(130, 105)
    scenes["cornell-c0"] = scene("cornell-c0", c0, materials(),
                                 comment="256x256, depth 4, 16 samples, 4 shadow rays")
    c1 = walls() + ceiling_light(130, 105) + [
        {"type": "sphere", "center": v(185.0, 90.0, -370.0), "radius": 90, "material": "mirror"},
        {"type": "sphere", "center": v(370.0, 90.0, -170.0), "radius": 90, "material": "glass"}]
    scenes["cornell-c1"] = scene("cornell-c1", c1, materials({"mirror": mirror, "glass": glass}),
                                 comment="720x720, depth 8, 64 spp, 4 shadow rays")
    c2 = walls() + ceiling_light(130, 105) + [
        {"type": "icosphere", "subdiv": 6, "seed": 42, "radius": 100.0, "displacement": 0.15,
         "center": v(L / 2, 130.0, -L / 2), "material": "white"}]
    scenes["cornell-c2"] = scene("cornell-c2", c2, materials(),
                                 comment="1024x1024, depth 8, 256 spp, 4 shadow rays")
    inside = (L / 2, L / 2, -12.0)
    c3 = walls(closed=True) + ceiling_light(30, 30) + [
        {"type": "sphere", "center": v(277.5, 90.0, -300.0), "radius": 90, "material": "white"}]
    scenes["closed-c3"] = scene("closed-c3", c3, materials(light_scale=400.0), cam=inside, fov=55.0,
                                comment="serving-stream scene")
    c4 = walls(closed=True) + ceiling_light(30, 30) + [
        {"type": "sphere", "center": v(277.5, 90.0, -300.0), "radius": 90, "material": "glass"}]
    scenes["closed-c4"] = scene("closed-c4", c4, materials({"glass": glass}, light_scale=400.0),
                                cam=inside, fov=55.0, comment="1024x1024, depth 8, hard indirect")
    for name, s in scenes.items():
        with open(os.path.join(HERE, name + ".json"), "w") as f:
            json.dump(s, f, indent=1)
            f.write("\n")


if __name__ == "__main__":
    main()
