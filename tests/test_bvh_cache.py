"""The content-keyed BVH cache of scene loads: a load served from memory or
from the cache directory gives the same nodes and triangle order as a fresh
build (the bit-exact restatement of the reference builder), and a damaged
cache file is rebuilt rather than trusted."""
import os

import numpy as np

import pcr_amd

HERE = os.path.dirname(os.path.abspath(__file__))
SC = os.path.join(HERE, "..", "pcr_amd", "scenes")


def load(name):
    h = pcr_amd.HostScene(os.path.join(SC, name + ".json"))
    return h.nodes(), h.tri_orig()


def test_cache_memory_disk_and_damage(tmp_path):
    pcr_amd.bvh_cache_dir(str(tmp_path))
    try:
        pcr_amd.bvh_cache_clear_memory()
        b0, m0, d0 = pcr_amd.bvh_cache_stats()
        fresh = load("cornell-c2")  # 81,932 triangles
        assert pcr_amd.bvh_cache_stats() == (b0 + 1, m0, d0)
        files = list(tmp_path.glob("*.bvh"))
        assert len(files) == 1
        mem = load("cornell-c2")
        assert pcr_amd.bvh_cache_stats() == (b0 + 1, m0 + 1, d0)
        pcr_amd.bvh_cache_clear_memory()
        disk = load("cornell-c2")
        assert pcr_amd.bvh_cache_stats() == (b0 + 1, m0 + 1, d0 + 1)
        for got in (mem, disk):
            assert np.array_equal(got[0], fresh[0]) and np.array_equal(got[1], fresh[1])
        # a damaged file (a leaf order that is no permutation) is rebuilt
        raw = bytearray(files[0].read_bytes())
        raw[-4:] = raw[-8:-4]
        files[0].write_bytes(bytes(raw))
        pcr_amd.bvh_cache_clear_memory()
        again = load("cornell-c2")
        assert pcr_amd.bvh_cache_stats()[0] == b0 + 2
        assert np.array_equal(again[0], fresh[0]) and np.array_equal(again[1], fresh[1])
        # other geometry gets its own entry
        load("cornell-c0")
        assert len(list(tmp_path.glob("*.bvh"))) == 2
    finally:
        pcr_amd.bvh_cache_dir(None)
        pcr_amd.bvh_cache_clear_memory()
