"""Reads scene files in the reference's JSON schema 1.0 (JSONC) for the tests.

Restates the ingestion rules of code/Scenes/SceneLoader.cpp:89-516 that
decide the numbers the renderer sees:
  - materials are indexed in sorted-key order (nlohmann's std::map objects),
  - numbers are parsed as double and narrowed to float (json get<float>),
  - SPD files are parsed with strtof and resampled (SpdLoader.cpp:38-80),
  - an SPD-only material derives its RGB from the integrated XYZ,
  - plane lights go only to the light list, triangle lights stay geometry.
The "icosphere" primitive is this project's procedural mesh.
"""
import ctypes as C
import json
import os

import numpy as np

from oracle_lib import OracleScene, f32, fp, lib, make_icosphere

_libc = C.CDLL(None)
_libc.strtof.restype = C.c_float
_libc.strtof.argtypes = [C.c_char_p, C.POINTER(C.c_char_p)]


def strip_jsonc(text):
    out = []
    i, n = 0, len(text)
    in_str = False
    while i < n:
        c = text[i]
        if in_str:
            out.append(c)
            if c == "\\" and i + 1 < n:
                out.append(text[i + 1])
                i += 2
                continue
            if c == '"':
                in_str = False
            i += 1
            continue
        if c == '"':
            in_str = True
            out.append(c)
            i += 1
        elif text.startswith("//", i):
            j = text.find("\n", i)
            i = n if j < 0 else j
        elif text.startswith("/*", i):
            j = text.find("*/", i + 2)
            i = n if j < 0 else j + 2
        else:
            out.append(c)
            i += 1
    return "".join(out)


def f(x):
    return np.float32(float(x))


def v3(a):
    assert isinstance(a, list) and len(a) == 3, a
    return np.array([f(x) for x in a], np.float32)


def load_spd(path):
    lam, val = [], []
    with open(path) as fh:
        for line in fh:
            s = line.lstrip(" \t")
            if not s.strip() or s.startswith("#"):
                continue
            parts = s.split()
            lam.append(_libc.strtof(parts[0].encode(), None))
            val.append(_libc.strtof(parts[1].encode(), None))
    lam = np.array(lam, np.float32)
    val = np.array(val, np.float32)
    out = np.zeros(61, np.float32)
    rc = lib().orc_spd_resample(fp(lam), fp(val), len(lam), fp(out))
    assert rc == 0, path
    return out


def resolve_spd(name, scene_path):
    d = os.path.dirname(os.path.abspath(scene_path))
    for root in (os.path.join(d, "spd"), os.path.join(d, "..", "spd"), os.path.join(os.getcwd(), "spd")):
        p = os.path.join(root, name + ".spd")
        if os.path.exists(p):
            return p
    raise FileNotFoundError("SPD %r not found for %s" % (name, scene_path))


def spectrum_rgb(tab):
    xyz = np.zeros(3, np.float32)
    rgb = np.zeros(3, np.float32)
    lib().orc_spectrum_to_xyz(fp(f32(tab)), fp(xyz))
    lib().orc_xyz_to_srgb(fp(xyz), fp(rgb))
    return np.maximum(rgb, np.float32(0))


class Material:
    def __init__(self):
        self.albedo = np.zeros(3, np.float32)
        self.emissive = np.zeros(3, np.float32)
        self.tab_albedo = None
        self.tab_emissive = None
        self.metallic = False
        self.transparent = False
        self.ior = np.float32(1.5)
        self.cauchyB = np.float32(0)


class SceneDesc:
    """Parsed scene in insertion order (before any BVH permutation)."""

    def __init__(self):
        self.name = ""
        self.version = ""
        self.camera = np.zeros(3, np.float32)
        self.fov = np.float32(65)
        self.materials = []
        self.mat_names = []
        self.spheres = []      # (center, radius, mat)
        self.walls = []        # (o, u, v, mat)
        self.plane_lights = []
        self.triangles = []    # (v[9], n[9] or None, mat)
        self.lights = []       # ("plane", idx) | ("tris", first, count) in declaration order
        self.ops = []          # primitives in declaration order (see build_oracle_scene)


def parse_scene(path):
    with open(path) as fh:
        j = json.loads(strip_jsonc(fh.read()))
    s = SceneDesc()
    s.name, s.version = j["name"], j["version"]
    cam = j.get("camera", {})
    if "position" in cam:
        s.camera = v3(cam["position"])
    if "fov" in cam:
        s.fov = f(cam["fov"])
    names = sorted(j["materials"].keys())
    idx = {}
    for n in names:
        m = j["materials"][n]
        mat = Material()
        if "albedo" in m:
            mat.albedo = v3(m["albedo"])
        if "emissive" in m:
            mat.emissive = v3(m["emissive"])
        if "albedo_spd" in m:
            mat.tab_albedo = load_spd(resolve_spd(m["albedo_spd"], path))
            if "albedo" not in m:
                mat.albedo = spectrum_rgb(mat.tab_albedo)
        if "emissive_spd" in m:
            tab = load_spd(resolve_spd(m["emissive_spd"], path))
            if "emissive_scale" in m:
                tab = (tab * f(m["emissive_scale"])).astype(np.float32)
            mat.tab_emissive = tab
            if "emissive" not in m:
                mat.emissive = spectrum_rgb(tab)
        mat.metallic = bool(m.get("metallic", False))
        mat.transparent = bool(m.get("transparent", False))
        if "ior" in m:
            mat.ior = f(m["ior"])
        if "cauchyB" in m:
            mat.cauchyB = f(m["cauchyB"])
        idx[n] = len(s.materials)
        s.materials.append(mat)
        s.mat_names.append(n)
    for p in j["primitives"]:
        t = p["type"]
        light = bool(p.get("light", False))
        if t == "sphere":
            s.spheres.append((v3(p["center"]), f(p["radius"]), idx[p["material"]]))
            s.ops.append(("sphere",) + s.spheres[-1])
        elif t == "triangle":
            v = np.concatenate([v3(p["v0"]), v3(p["v1"]), v3(p["v2"])])
            nrm = None
            if "n0" in p:
                nrm = np.concatenate([v3(p["n0"]), v3(p["n1"]), v3(p["n2"])])
            if light:
                s.lights.append(("tris", len(s.triangles), 1))
            s.triangles.append((v, nrm, idx[p["material"]]))
            s.ops.append(("tri", v, nrm, idx[p["material"]], light))
        elif t == "plane":
            pl = (v3(p["origin"]), v3(p["u"]), v3(p["v"]), idx[p["material"]])
            if light:
                s.lights.append(("plane", len(s.plane_lights)))
                s.plane_lights.append(pl)
            else:
                s.walls.append(pl)
            s.ops.append(("plane",) + pl + (light,))
        elif t == "icosphere":
            tris = make_icosphere(int(p["subdiv"]), int(p["seed"]), float(p["radius"]),
                                  float(p["displacement"]), [float(x) for x in v3(p["center"])])
            first = len(s.triangles)
            m = idx[p["material"]]
            for tri in tris:
                s.triangles.append((tri, None, m))
            if light:
                s.lights.append(("tris", first, len(tris)))
            s.ops.append(("mesh", first, len(tris), m, light, tris))
        else:
            raise ValueError("unsupported primitive type %r" % t)
    return s


def triangle_array(desc):
    return np.array([t[0] for t in desc.triangles], np.float32).reshape(-1, 9)


def build_oracle_scene(desc):
    """Replays the primitives in declaration order, so the light list, the
    wall list and the triangle list come out in the reference's order."""
    sc = OracleScene()
    L = lib()
    for m in desc.materials:
        ta = fp(f32(m.tab_albedo)) if m.tab_albedo is not None else None
        te = fp(f32(m.tab_emissive)) if m.tab_emissive is not None else None
        L.orc_scene_add_material(sc.h, fp(f32(m.albedo)), fp(f32(m.emissive)), int(m.metallic),
                                 int(m.transparent), float(m.ior), float(m.cauchyB), ta, te)
    for op in desc.ops:
        kind = op[0]
        if kind == "sphere":
            _, c, r, m = op
            L.orc_scene_add_sphere(sc.h, fp(f32(c)), float(r), m)
        elif kind == "plane":
            _, o, u, v, m, light = op
            L.orc_scene_add_plane(sc.h, fp(f32(o)), fp(f32(u)), fp(f32(v)), m, int(light))
        elif kind == "tri":
            _, v, nrm, m, light = op
            L.orc_scene_add_triangle(sc.h, fp(f32(v)), fp(f32(nrm)) if nrm is not None else None, m,
                                     int(light))
        elif kind == "mesh":
            _, first, count, m, light, tris = op
            for tri in tris:
                L.orc_scene_add_triangle(sc.h, fp(f32(tri)), None, m, 0)
            if light:
                L.orc_scene_add_mesh_light(sc.h, first, count)
    L.orc_scene_set_camera(sc.h, fp(f32(desc.camera)), float(desc.fov))
    L.orc_scene_finalize(sc.h)
    sc.n_tris = len(desc.triangles)
    return sc
