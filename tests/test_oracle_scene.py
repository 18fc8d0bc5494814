"""Oracle scene ingestion + sceneIntersect vs the compiled reference's
loadSceneFromFile (SceneLoader.cpp:475-516) and Renderer::sceneIntersect
(Renderer.cpp:891-949), on 4096 seeded rays per scene."""
import os

import numpy as np
import pytest

import scene_io

HERE = os.path.dirname(os.path.abspath(__file__))
REF = ["cornell", "cornell-glass", "cornell-spec", "cornell-large-light", "cornell-spheres",
       "closed-cornell-glass", "closed-cornell-spec", "bright-cornell-glass"]
OWN = ["cornell-c0", "cornell-c1", "closed-c3", "closed-c4"]
CASES = [("ref_" + n, os.path.join(HERE, "golden", "ref", "Scenes", n + ".json")) for n in REF] + \
        [(n, os.path.join(HERE, "..", "pcr_amd", "scenes", n + ".json")) for n in OWN]


def u32(a):
    return np.asarray(a, np.float32).view(np.uint32)


@pytest.mark.parametrize("key,path", CASES, ids=[c[0] for c in CASES])
def test_scene_matches_reference(key, path):
    g = np.load(os.path.join(HERE, "golden", "scene_%s.npz" % key))
    desc = scene_io.parse_scene(path)
    sc = scene_io.build_oracle_scene(desc)
    fa, fe = sc.fits()
    assert np.array_equal(u32(fa), u32(g["fit_albedo"]))
    assert np.array_equal(u32(fe), u32(g["fit_emissive"]))
    assert np.array_equal(u32([m.albedo for m in desc.materials]), u32(g["albedo"]))
    assert np.array_equal(u32([m.emissive for m in desc.materials]), u32(g["emissive"]))
    nodes, perm = sc.bvh()
    assert np.array_equal(nodes, g["nodes"])
    tris = scene_io.triangle_array(desc)
    if len(tris):
        assert np.array_equal(u32(tris[perm]), u32(g["leaf_tris"]))
    info = sc.info()
    assert u32(info.total_light_area) == u32(g["total_light_area"])[0]
    h = sc.intersect(g["rays"])
    kind = g["hit_kind"]
    assert np.array_equal(h["kind"], kind)
    assert np.array_equal(h["mat"], g["hit_mat"])
    ok = kind >= 0
    assert np.array_equal(u32(h["t"][ok]), u32(g["hit_t"][ok]))
    assert np.array_equal(u32(h["hit"][ok]), u32(g["hit_point"][ok]))
    assert np.array_equal(u32(h["normal"][ok]), u32(g["hit_normal"][ok]))
    ref_prim = g["hit_prim"].copy()
    tri = kind == 2
    ref_prim[tri] = perm[ref_prim[tri]]  # reference reports leaf-order triangle indices
    assert np.array_equal(h["prim"][ok], ref_prim[ok])
