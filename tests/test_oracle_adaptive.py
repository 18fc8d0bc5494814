"""Pins of the oracle's adaptive sampling with the per-path estimator
(orc_render_paths_adaptive; the rule of Renderer.cpp:266-294, a point being
the mean of one AA iteration of `group` paths as in MetalRenderer.mm:1500-1540):

  * it equals a float32 restatement of the rule driven by the per-group sums
    of the (already pinned) non-adaptive per-path renderer, sample ranges
    [g * group, (g + 1) * group): the same Welford points, the same stop, the
    same number of paths traced;
  * with one group per pixel (group = spp) no pixel can stop (a stop needs 4
    points), and the mean is the plain mean sum / spp rounded to float.
"""
import os

import numpy as np
import pytest

import scene_io

HERE = os.path.dirname(os.path.abspath(__file__))
SC = os.path.join(HERE, "..", "pcr_amd", "scenes")


def oracle(name):
    return scene_io.build_oracle_scene(scene_io.parse_scene(os.path.join(SC, name + ".json")))


def welford_reference(groups, group):
    """groups: (G, h, w, 3) FP64 sums of each group.  float32 Welford with the
    stop rule; returns means (h, w, 3) float32 and the paths traced."""
    G, h, w, _ = groups.shape
    out = np.zeros((h, w, 3), np.float32)
    traced = 0
    f = np.float32
    for y in range(h):
        for x in range(w):
            mean = [f(0), f(0), f(0)]
            m2 = [f(0), f(0), f(0)]
            points = 0
            for g in range(G):
                traced += group
                points += 1
                conv = points >= 4
                for c in range(3):
                    m = f(groups[g, y, x, c] / group)
                    d = f(m - mean[c])
                    mean[c] = f(mean[c] + f(d / f(points)))
                    m2[c] = f(m2[c] + f(d * f(m - mean[c])))
                    if points >= 4:
                        v = f(m2[c] / f(points - 1))
                        conv = conv and f(v / f(f(mean[c] * mean[c]) + f(0.01))) < f(0.05)
                if conv:
                    break
            out[y, x] = mean
    return out, traced


@pytest.mark.parametrize("name,spectral", [("cornell-c1", False), ("cornell-c1", True), ("closed-c4", True)])
def test_adaptive_matches_restatement(name, spectral):
    sc = oracle(name)
    W = H = 12
    group, G = 3, 8
    kw = dict(width=W, height=H, depth=5, spp=group * G, shadow=2, seed=4321, jitter=True, spectral=spectral)
    groups = np.stack([sc.render_paths(**kw, s_range=(g * group, (g + 1) * group))[0] for g in range(G)])
    want, traced = welford_reference(groups, group)
    got, cnt = sc.render_paths_adaptive(**kw, group=group)
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))
    assert int(cnt[0]) == traced
    assert traced < W * H * group * G  # some pixels stopped early


def test_adaptive_single_group_is_plain_mean():
    sc = oracle("cornell-c1")
    kw = dict(width=10, height=10, depth=5, spp=6, shadow=2, seed=99, jitter=True)
    sums, _ = sc.render_paths(**kw)
    got, cnt = sc.render_paths_adaptive(**kw, group=6)
    assert np.array_equal(got, (sums / 6).astype(np.float32))
    assert int(cnt[0]) == 10 * 10 * 6
