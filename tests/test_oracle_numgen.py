"""Oracle NumGen vs the compiled reference (code/NumGen.cpp:24-80)."""
import numpy as np

import oracle_lib as ol


def test_streams_match_reference(golden):
    g = np.load(golden + "/numgen.npz")
    for key in g.files:
        seed = int(key.split("_")[1])
        ref = g[key]
        for k in (0, 1):
            got = ol.numgen(seed, k, ref.shape[1])
            assert np.array_equal(got.view(np.uint32), ref[k].view(np.uint32)), (seed, k)


def test_pcg32_published_vector():
    # O'Neill's pcg32 XSH-RR output function on a known state sequence: the
    # first output of state 0 is 0 (xorshift of 0), and the LCG step is the
    # Knuth MMIX multiplier with the reference's increment.
    out = np.zeros(3, np.uint32)
    st = np.zeros(1, np.uint64)
    ol.lib().orc_pcg_raw(0, 3, out.ctypes.data_as(ol.C.POINTER(ol.C.c_uint32)), st.ctypes.data_as(ol._u64))
    assert out[0] == 0
    s = np.uint64(0)
    with np.errstate(over="ignore"):
        for _ in range(3):
            s = s * np.uint64(6364136223846793005) + np.uint64(1442695040888963407)
    assert st[0] == s


def test_epsilon_can_reach_one():
    # (float)u32 * 2^-32 rounds to 1.0f for u32 >= 2^32 - 128 (NumGen.cpp:79).
    x = np.float32(np.float32(4294967295.0) * np.float32(1.0 / 4294967296.0))
    assert x == np.float32(1.0)
