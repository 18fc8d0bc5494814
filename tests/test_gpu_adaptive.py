"""Adaptive sampling on the GPU (PCRHIP_ADAPTIVE): k_resolve closes each
group of samples into a float Welford point and marks converged pixels,
whose later paths are not traced.  The Welford means must equal the oracle's
restatement (orc_render_paths_adaptive) bit for bit, with groups straddling
pass boundaries; with one sample per pass the GPU also traces exactly the
oracle's number of paths, and with larger passes at least as many (a pass
finishes the samples it started) but still fewer than all."""
import os

import numpy as np
import pytest

import pcr_amd
import scene_io

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
SC = os.path.join(HERE, "..", "pcr_amd", "scenes")


@pytest.fixture
def ctx_cfg(gpu_ctx):
    saved = gpu_ctx.config()
    yield gpu_ctx
    gpu_ctx.set_config(**{f: getattr(saved, f) for f, _ in pcr_amd.Config._fields_})


@pytest.mark.parametrize("name,spectral", [("cornell-c1", False), ("cornell-c1", True), ("closed-c4", True)])
def test_adaptive_bit_exact(ctx_cfg, name, spectral):
    path = os.path.join(SC, name + ".json")
    W = H = 32  # one sample per pass is 1024 paths, the smallest pass
    group, G = 3, 8
    kw = dict(width=W, height=H, depth=5, spp=group * G, shadow=2, jitter=True, spectral=spectral)
    want, cnt = scene_io.build_oracle_scene(scene_io.parse_scene(path)).render_paths_adaptive(
        seed=4321, group=group, **kw)
    ds = ctx_cfg.upload(pcr_amd.HostScene(path))
    try:
        for pass_spp, exact in ((1, True), (5, False)):
            ctx_cfg.set_config(target_pass_paths=W * H * pass_spp)
            j = ds.submit(seed=4321, adaptive_group=group, **kw)
            j.wait()
            got = j.sums().astype(np.float32)
            img, st = j.result()
            j.release()
            assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), pass_spp
            assert np.array_equal(img, want)
            if exact:
                assert st.paths == int(cnt[0])
            assert int(cnt[0]) <= st.paths < W * H * group * G, (st.paths, int(cnt[0]))
    finally:
        ds.release()


def test_adaptive_rejects_bad_group(gpu_ctx):
    ds = gpu_ctx.upload(pcr_amd.HostScene(os.path.join(SC, "cornell-c1.json")))
    try:
        with pytest.raises(pcr_amd.PcrError):
            ds.submit(width=8, height=8, depth=2, spp=10, shadow=1, seed=1, adaptive_group=3)
    finally:
        ds.release()
