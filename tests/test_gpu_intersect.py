"""HIP closest-hit kernel vs the reference's sceneIntersect / Bvh::intersect:
hit kind, primitive (original triangle index), material, t, hit point and
normal, bit for bit, on the reference-generated ray sets of 12 scenes and the
81,932-triangle displaced icosphere, plus 200k seeded rays vs the oracle."""
import os

import numpy as np
import pytest

import pcr_amd
import scene_io

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
G = os.path.join(HERE, "golden")
REF = ["cornell", "cornell-glass", "cornell-spec", "cornell-large-light", "cornell-spheres",
       "closed-cornell-glass", "closed-cornell-spec", "bright-cornell-glass"]
OWN = ["cornell-c0", "cornell-c1", "closed-c3", "closed-c4"]
CASES = [("ref_" + n, os.path.join(G, "ref", "Scenes", n + ".json")) for n in REF] + \
        [(n, os.path.join(HERE, "..", "pcr_amd", "scenes", n + ".json")) for n in OWN]


def u32(a):
    return np.asarray(a, np.float32).view(np.uint32)


def compare(h, t, kind, prim, mat, hit=None, normal=None):
    assert np.array_equal(h["kind"], kind)
    ok = kind >= 0
    assert np.array_equal(h["prim"][ok], prim[ok])
    assert np.array_equal(h["mat"][ok], mat[ok])
    assert np.array_equal(u32(h["t"][ok]), u32(t[ok]))
    if hit is not None:
        assert np.array_equal(u32(h["hit"][ok]), u32(hit[ok]))
    if normal is not None:
        assert np.array_equal(u32(h["normal"][ok]), u32(normal[ok]))


@pytest.mark.parametrize("key,path", CASES, ids=[c[0] for c in CASES])
def test_fixture_rays(gpu_ctx, key, path):
    print("intersect", key, flush=True)
    g = np.load(os.path.join(G, "scene_%s.npz" % key))
    hs = pcr_amd.HostScene(path)
    ds = gpu_ctx.upload(hs)
    try:
        h = ds.intersect(g["rays"])
    finally:
        ds.release()
    prim = g["hit_prim"].copy()
    tri = g["hit_kind"] == 2
    prim[tri] = hs.tri_orig()[prim[tri]]  # reference reports leaf-order indices
    compare(h, g["hit_t"], g["hit_kind"], prim, g["hit_mat"], g["hit_point"], g["hit_normal"])


def test_displaced_icosphere_rays(gpu_ctx, scenes_dir):
    print("intersect cornell-c2", flush=True)
    g = np.load(os.path.join(G, "bvh_c2.npz"))
    hs = pcr_amd.HostScene(os.path.join(scenes_dir, "cornell-c2.json"))
    ds = gpu_ctx.upload(hs)
    try:
        info = ds.info()
        assert info.bvh_depth == int(g["stats"][0])
        h = ds.intersect(g["rays"])
    finally:
        ds.release()
    idx = g["hit_index"]
    assert np.array_equal(np.where(h["kind"] == 2, h["prim"], -1), idx)
    ok = idx >= 0
    assert np.array_equal(u32(h["t"][ok]), u32(g["hit_t"][ok]))
    assert np.array_equal(u32(h["hit"][ok]), u32(g["hit_point"][ok]))
    assert np.array_equal(u32(h["normal"][ok]), u32(g["hit_normal"][ok]))


@pytest.mark.parametrize("name", ["cornell-c1", "cornell-c2"])
def test_many_rays_vs_oracle(gpu_ctx, scenes_dir, name):
    print("intersect 200k", name, flush=True)
    path = os.path.join(scenes_dir, name + ".json")
    rng = np.random.default_rng(99)
    n = 200_000
    o = rng.uniform([1, 1, -554], [554, 554, -1], size=(n, 3))
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rays = np.concatenate([o, d], 1).astype(np.float32)
    oracle = scene_io.build_oracle_scene(scene_io.parse_scene(path)).intersect(rays)
    ds = gpu_ctx.upload(pcr_amd.HostScene(path))
    try:
        h = ds.intersect(rays)
    finally:
        ds.release()
    compare(h, oracle["t"], oracle["kind"], oracle["prim"], oracle["mat"], oracle["hit"], oracle["normal"])
