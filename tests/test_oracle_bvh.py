"""Oracle BVH build/traversal vs the compiled reference (code/Bvh/Bvh.cpp) and
the reference's own tests/unit/bvh.cpp cases."""
import hashlib
import os

import numpy as np
import pytest

import oracle_lib as ol
import scene_io


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def tri_scene(tris):
    sc = ol.OracleScene()
    L = ol.lib()
    L.orc_scene_add_material(sc.h, ol.fp(ol.f32([0.5, 0.5, 0.5])), ol.fp(ol.f32([0, 0, 0])), 0, 0, 1.5, 0, None, None)
    L.orc_scene_add_material(sc.h, ol.fp(ol.f32([0, 0, 0])), ol.fp(ol.f32([1, 1, 1])), 0, 0, 1.5, 0, None, None)
    for t in ol.f32(tris).reshape(-1, 9):
        L.orc_scene_add_triangle(sc.h, ol.fp(t.copy()), None, 0, 0)
    # a light far away so the scene is valid; it is a plane, not in the BVH
    L.orc_scene_add_plane(sc.h, ol.fp(ol.f32([1e6, 1e6, 1e6])), ol.fp(ol.f32([1, 0, 0])),
                          ol.fp(ol.f32([0, 1, 0])), 1, 1)
    L.orc_scene_finalize(sc.h)
    return sc


def test_mt19937_set_bitexact(golden):
    g = np.load(golden + "/bvh_mt19937.npz")
    nodes, perm = ol.bvh_build(g["tris"])
    assert np.array_equal(nodes, g["nodes"])
    assert np.array_equal(perm, g["perm"])
    sc = tri_scene(g["tris"])
    n2, p2 = sc.bvh()
    assert np.array_equal(n2, g["nodes"]) and np.array_equal(p2, g["perm"])
    info = sc.info()
    assert [info.bvh_depth, info.bvh_leaves, info.bvh_max_leaf] == list(g["stats"])
    assert info.bvh_depth < 30  # bvh.cpp:109-121
    hits = sc.intersect(g["rays"], tmax=1e30)
    got_idx = np.where(hits["kind"] == 2, hits["prim"], -1)
    assert np.array_equal(got_idx, g["hit_index"])
    ok = g["hit_index"] >= 0
    assert np.array_equal(hits["t"][ok].view(np.uint32), g["hit_t"][ok].view(np.uint32))


def brute(tris, rays):
    """Brute-force closest hit (bvh.cpp:125-139) through the oracle with no BVH
    pruning: one scene per triangle would be slow, so test each ray against
    each triangle scene-free via a 1-triangle BVH."""
    best_t = np.full(len(rays), np.inf, np.float32)
    best_i = np.full(len(rays), -1, np.int32)
    for i, t in enumerate(ol.f32(tris).reshape(-1, 9)):
        sc = tri_scene(t[None])
        h = sc.intersect(rays, tmax=1e30)
        hit = (h["kind"] == 2) & (h["t"] < best_t)
        best_t[hit] = h["t"][hit]
        best_i[hit] = i
    return best_i, best_t


def test_bvh_equals_brute_force(golden):
    g = np.load(golden + "/bvh_mt19937.npz")
    rng = np.random.default_rng(3)
    o = np.stack([rng.uniform(-1, 1, 300), rng.uniform(-1, 1, 300), np.zeros(300)], 1)
    d = np.stack([rng.uniform(-0.2, 0.2, 300), rng.uniform(-0.2, 0.2, 300), -np.ones(300)], 1)
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rays = ol.f32(np.concatenate([o, d], 1))
    bi, bt = brute(g["tris"], rays)
    h = tri_scene(g["tris"]).intersect(rays, tmax=1e30)
    got = np.where(h["kind"] == 2, h["prim"], -1)
    assert np.array_equal(got, bi)
    assert np.array_equal(h["t"][bi >= 0], bt[bi >= 0])


def test_degenerate_builds():
    nodes, perm = ol.bvh_build(np.zeros((0, 9), np.float32))
    assert len(nodes) == 0
    nodes, perm = ol.bvh_build(ol.f32([[-1, -1, -5, 1, -1, -5, 0, 1, -5]]))
    assert len(nodes) == 1 and nodes[0, 8] == 1
    cop = []
    for i in range(100):
        x = i * 0.01
        cop.append([x, 0, -5, x + 0.005, 0, -5, x, 0.01, -5])
    sc = tri_scene(np.array(cop, np.float32))
    assert sc.info().bvh_depth < 60  # bvh.cpp:71-89


def test_c0_bitexact(golden, scenes_dir):
    g = np.load(golden + "/bvh_c0.npz")
    desc = scene_io.parse_scene(os.path.join(scenes_dir, "cornell-c0.json"))
    tris = scene_io.triangle_array(desc)
    assert sha(tris) == str(g["tris_sha256"])
    nodes, perm = ol.bvh_build(tris)
    assert np.array_equal(nodes, g["nodes"]) and np.array_equal(perm, g["perm"])


def test_c2_bitexact(golden, scenes_dir):
    g = np.load(golden + "/bvh_c2.npz")
    desc = scene_io.parse_scene(os.path.join(scenes_dir, "cornell-c2.json"))
    tris = scene_io.triangle_array(desc)
    assert len(tris) == 81932
    assert sha(tris) == str(g["tris_sha256"]), "procedural mesh changed"
    nodes, perm = ol.bvh_build(tris)
    assert len(nodes) == int(g["n_nodes"])
    assert sha(nodes) == str(g["nodes_sha256"]) and sha(perm) == str(g["perm_sha256"])
    # hits of the raw-triangle BVH on 4096 rays
    sc = tri_scene(tris)
    h = sc.intersect(g["rays"])
    got = np.where(h["kind"] == 2, h["prim"], -1)
    assert np.array_equal(got, g["hit_index"])
    ok = got >= 0
    assert np.array_equal(h["t"][ok].view(np.uint32), g["hit_t"][ok].view(np.uint32))
    assert np.array_equal(h["hit"][ok].view(np.uint32), g["hit_point"][ok].view(np.uint32))
    assert np.array_equal(h["normal"][ok].view(np.uint32), g["hit_normal"][ok].view(np.uint32))
