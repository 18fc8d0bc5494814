"""Image-level statistics of HIP renders against the CPU.

For fixed crops of the 256x256 box scene (depth 4, 16 spp) and the 720x720
mirror + dispersive glass scene (depth 8, 64 spp), against 65536-sample FP64
CPU references (tests/golden/stats_ref.npz, made by make_stats_ref.py):

  * relMSE of the HIP renders <= 1.05 x the relMSE of CPU renders at equal
    spp, both averaged over 32 independent seeds (disjoint seed sets, so the
    two sides are independent samples, not the same paths);
  * per-channel crop means of the HIP renders within 3 standard errors of
    the CPU renders' (errors from the seed-to-seed spread).
"""
import math
import os

import numpy as np
import pytest

import oracle_lib as ol
import pcr_amd
import scene_io

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
SC = os.path.join(HERE, "..", "pcr_amd", "scenes")
SCENE = {"c0": ("cornell-c0", 16), "c1": ("cornell-c1", 64)}
K = 32
EPS = 1e-3  # relMSE denominator floor (XYZ units; the crops' values are 1..100)


def relmse(img, ref):
    return float(np.mean((img - ref) ** 2 / (ref ** 2 + EPS)))


@pytest.mark.parametrize("key", sorted(SCENE))
def test_relmse_and_means_vs_cpu(gpu_ctx, golden, key):
    g = np.load(os.path.join(golden, "stats_ref.npz"))
    ref = g[key + "_mean"]
    W, H, d, S, x0, y0, cw, ch, _, _ = (int(v) for v in g[key + "_meta"])
    name, spp = SCENE[key]
    path = os.path.join(SC, name + ".json")
    osc = scene_io.build_oracle_scene(scene_io.parse_scene(path))
    ds = gpu_ctx.upload(pcr_amd.HostScene(path))
    try:
        gpu, cpu = [], []
        for k in range(K):
            j = ds.submit(width=W, height=H, depth=d, spp=spp, shadow=S, seed=ol.spread_seed(100 + k), crop=(x0, y0, cw, ch))
            j.wait()
            gpu.append(j.sums() / spp)
            j.release()
            acc, _ = osc.render_paths(W, H, d, spp, S, ol.spread_seed(900 + k), crop=(x0, y0, cw, ch))
            cpu.append(acc / spp)
    finally:
        ds.release()
    r_gpu = np.mean([relmse(x, ref) for x in gpu])
    r_cpu = np.mean([relmse(x, ref) for x in cpu])
    print("%s: relMSE gpu %.5g cpu %.5g ratio %.4f" % (key, r_gpu, r_cpu, r_gpu / r_cpu))
    assert r_gpu <= 1.05 * r_cpu
    mg = np.array([x.reshape(-1, 3).mean(0) for x in gpu])
    mc = np.array([x.reshape(-1, 3).mean(0) for x in cpu])
    se = np.sqrt(mg.var(0, ddof=1) / K + mc.var(0, ddof=1) / K)
    z = (mg.mean(0) - mc.mean(0)) / se
    assert np.all(np.abs(z) < 3.0), z
    # and both agree with the 65536-sample reference's crop mean
    zr = (mg.mean(0) - ref.reshape(-1, 3).mean(0)) / (mg.std(0, ddof=1) / math.sqrt(K))
    assert np.all(np.abs(zr) < 3.0), zr
