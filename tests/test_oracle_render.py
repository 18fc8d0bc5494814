"""Oracle branched estimator vs the reference's seeded golden PNGs
(tests/render/golden, made by tests/render/run.sh with -d 2 -s 4 -S 2 -w 120
--seed 12345) and vs reference CLI renders of the benchmark scenes
(tests/golden/renders).  Byte-identical RGB8 is required."""
import os

import numpy as np
import pytest
from PIL import Image

import scene_io

HERE = os.path.dirname(os.path.abspath(__file__))
G = os.path.join(HERE, "golden")

# key -> (scene, flags) from tests/render/run.sh:49-67 (the reproducible subset:
# no LUT, no denoiser, no OBJ mesh)
REF_TUPLES = {
    "cornell-rgb": ("cornell", {}),
    "cornell-spectral": ("cornell", {"spectral": True}),
    "cornell-spectral-hero1": ("cornell", {"spectral": True, "hero": 1}),
    "cornell-glass-rgb": ("cornell-glass", {}),
    "cornell-glass-spectral": ("cornell-glass", {"spectral": True}),
    "cornell-spheres-rgb": ("cornell-spheres", {}),
    "cornell-large-light-rgb": ("cornell-large-light", {}),
    "cornell-spec-rgb": ("cornell-spec", {}),
    "cornell-rgb-aces": ("cornell", {"aces": True}),
    "cornell-rgb-mis-russian-stratified": ("cornell", {"mis": True, "russian": True, "stratified": True}),
    "cornell-rgb-aa4": ("cornell", {"aa": 4}),
    "cornell-rgb-adaptive-aa8": ("cornell", {"adaptive": True, "aa": 8}),
}


def load_png(path):
    return np.asarray(Image.open(path).convert("RGB"))


@pytest.mark.parametrize("key", sorted(REF_TUPLES))
def test_reference_golden(key):
    scene, flags = REF_TUPLES[key]
    desc = scene_io.parse_scene(os.path.join(G, "ref", "Scenes", scene + ".json"))
    sc = scene_io.build_oracle_scene(desc)
    _, rgb8 = sc.render_branched(120, 120, 2, 4, 2, 12345, **flags)
    want = load_png(os.path.join(G, "ref", "golden", key + ".png"))
    diff = np.count_nonzero(np.any(rgb8 != want, axis=2))
    assert diff == 0, "%d pixels differ" % diff


def parse_flags(args):
    out = {"spectral": False}
    it = iter(args)
    kw = {}
    for a in it:
        if a == "-d": kw["depth"] = int(next(it))
        elif a == "-s": kw["samples"] = int(next(it))
        elif a == "-S": kw["shadow"] = int(next(it))
        elif a == "-w": kw["width"] = int(next(it))
        elif a == "--spectral": out["spectral"] = True
        elif a == "--mis": out["mis"] = True
        elif a == "--russian": out["russian"] = True
        elif a == "--stratified": out["stratified"] = True
        elif a == "--aa": out.setdefault("aa", 4)
        elif a == "--aa-samples": out["aa"] = int(next(it))
    return kw, out


def own_renders():
    rows = []
    with open(os.path.join(G, "renders", "README.txt")) as f:
        for line in f:
            if ": --scene " in line:
                key, rest = line.strip().split(": --scene ")
                parts = rest.split()
                rows.append((key, parts[0], parts[1:]))
    return rows


@pytest.mark.parametrize("key,scene,args", own_renders(), ids=[r[0] for r in own_renders()])
def test_benchmark_scene_render(key, scene, args):
    kw, flags = parse_flags(args)
    desc = scene_io.parse_scene(os.path.join(HERE, "..", "pcr_amd", "scenes", scene + ".json"))
    sc = scene_io.build_oracle_scene(desc)
    w = kw["width"]
    _, rgb8 = sc.render_branched(w, w, kw["depth"], kw["samples"], kw["shadow"], 12345, **flags)
    want = load_png(os.path.join(G, "renders", key + ".png"))
    diff = np.count_nonzero(np.any(rgb8 != want, axis=2))
    assert diff == 0, "%d pixels differ" % diff
