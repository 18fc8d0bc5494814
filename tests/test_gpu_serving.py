"""Serving loop: a small seeded Poisson stream of mixed jobs (4 scenes,
resolutions, spp and depths), run with continuous batching and one job at a
time.  Every job's FP64 result must be identical in both runs -- the same alone
and in the mix -- and equal to the CPU oracle's; latency percentiles and
throughput are reported."""
import os

import numpy as np
import pytest

import scene_io
from pcr_amd import serving

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
SC = os.path.join(HERE, "..", "pcr_amd", "scenes")


def test_stream_alone_equals_mixed():
    stream = serving.make_stream(n=24, rate=300.0, seed=7, resolutions=(24, 32, 48), spp_range=(4, 48))
    report, rb, ro = serving.benchmark(stream, keep_sums=True, progress=0, log=lambda *a: None)
    print(report)
    for spec, b, o in zip(stream, rb, ro):
        assert np.array_equal(b["sums"].view(np.uint64), o["sums"].view(np.uint64)), spec
    # three jobs against the oracle
    for i in (0, 5, 11):
        spec = stream[i]
        d = dict(spec["desc"])
        osc = scene_io.build_oracle_scene(scene_io.parse_scene(os.path.join(SC, serving.SCENES[spec["scene"]] + ".json")))
        want, _ = osc.render_paths(d.pop("width"), d.pop("height"), d.pop("depth"), d.pop("spp"), d.pop("shadow"),
                                   d.pop("seed"), **d)
        assert np.array_equal(rb[i]["sums"].view(np.uint64), want.view(np.uint64)), i
    for k in ("batched", "one_at_a_time"):
        r = report[k]
        assert r["p50_ms"] > 0 and r["p99_ms"] >= r["p50_ms"] and r["jobs_per_s"] > 0
