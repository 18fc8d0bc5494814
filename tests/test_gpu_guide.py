"""The learned proposal on the GPU:

  * the MLP forward pass (encoding + three BF16 MFMA layers, FP32
    accumulation) against the oracle's FP64 restatement with the same BF16
    roundings (differences come only from FP32 accumulation order and the
    BF16 rounding it can flip: 5e-3 of the output scale);
  * the training gradient (backward pass as MFMA tiles) against the oracle's
    backpropagation: relative L2 error < 2e-2, cosine > 0.999 -- the oracle's
    own gradient is checked against central differences on the CPU;
  * a guided render is unbiased: per-channel image means within 3 standard
    errors of the plain render's (cosine sampling keeps half the mixture),
    with the default 3 guided vertices per path and with every vertex guided;
  * training lowers the loss and the guided estimator beats plain sampling at
    equal samples on the closed box with the small light (relMSE against a
    high-sample plain render).
"""
import math
import os

import numpy as np
import pytest

import oracle_lib as ol
import pcr_amd

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
SC = os.path.join(HERE, "..", "pcr_amd", "scenes")
BOX = np.array([0, 0, -555, 555, 555, 0], np.float32)


def random_verts(g, n):
    v = np.zeros((n, 6), np.float32)
    v[:, 0:3] = g.uniform([0, 0, -555], [555, 555, 0], (n, 3))
    nrm = g.normal(size=(n, 3))
    v[:, 3:6] = nrm / np.linalg.norm(nrm, axis=1, keepdims=True)
    return v


def test_forward_matches_oracle(gpu_ctx):
    g = np.random.default_rng(1)
    for seed in (3, 4):
        params = ol.guide_random_params(seed)
        verts = random_verts(g, 300)
        got = gpu_ctx.guide_forward(params, BOX, verts)
        want = ol.guide_forward(params, BOX, verts, quantize=True)
        scale = max(1.0, float(np.abs(want).max()))
        assert np.max(np.abs(got - want)) <= 5e-3 * scale, np.max(np.abs(got - want))


def test_gradient_matches_oracle(gpu_ctx):
    g = np.random.default_rng(2)
    params = ol.guide_random_params(5)
    n = 300
    recs = np.zeros((n, 10), np.float32)
    recs[:, 0:6] = random_verts(g, n)
    d = g.normal(size=(n, 3))
    recs[:, 6:9] = d / np.linalg.norm(d, axis=1, keepdims=True)
    recs[:, 9] = g.uniform(0.0, 2.0, n)
    recs[::7, 9] = 0.0  # zero-weight records contribute nothing
    got, loss = gpu_ctx.guide_gradient(params, BOX, recs)
    want, wloss = ol.guide_gradient(params, BOX, recs, quantize=True)
    err = np.linalg.norm(got - want) / np.linalg.norm(want)
    cos = float(got @ want / (np.linalg.norm(got) * np.linalg.norm(want)))
    assert err < 2e-2 and cos > 0.999, (err, cos)
    assert abs(loss - wloss) <= 1e-3 * abs(wloss)


def render_mean(ds, seed, **kw):
    j = ds.submit(seed=seed, **kw)
    j.wait()
    s = j.sums()
    losses = j.guide_losses() if kw.get("guided") else None
    j.release()
    return s / kw["spp"], losses


@pytest.mark.parametrize("bounces", [0, 8])  # the default (3 guided vertices per path), every vertex
def test_guided_render_unbiased(gpu_ctx, bounces):
    ds = gpu_ctx.upload(pcr_amd.HostScene(os.path.join(SC, "closed-c4.json")))
    try:
        kw = dict(width=96, height=96, depth=8, spp=32, shadow=4)
        K = 8
        plain = np.array([render_mean(ds, ol.spread_seed(40 + k), **kw)[0].reshape(-1, 3).mean(0) for k in range(K)])
        guided = np.array([render_mean(ds, ol.spread_seed(80 + k), guided=True, guide_bounces=bounces,
                                       **kw)[0].reshape(-1, 3).mean(0) for k in range(K)])
        se = np.sqrt(plain.var(0, ddof=1) / K + guided.var(0, ddof=1) / K)
        z = (guided.mean(0) - plain.mean(0)) / se
        print("plain", plain.mean(0), "guided", guided.mean(0), "z", z)
        assert np.all(np.abs(z) < 3.0), z
    finally:
        ds.release()


def test_guided_training_gain(gpu_ctx):
    ds = gpu_ctx.upload(pcr_amd.HostScene(os.path.join(SC, "closed-c4.json")))
    try:
        kw = dict(width=128, height=128, depth=8, shadow=4)
        ref, _ = render_mean(ds, 12345, spp=4096, **kw)
        def relmse(x):
            return float(np.mean((x - ref) ** 2 / (ref ** 2 + 1e-2)))
        spp = 256
        rp = np.mean([relmse(render_mean(ds, ol.spread_seed(200 + k), spp=spp, **kw)[0]) for k in range(3)])
        rg, losses = [], []
        for k in range(3):
            img, ls = render_mean(ds, ol.spread_seed(300 + k), spp=spp, guided=True, **kw)
            rg.append(relmse(img))
            losses.append(ls)
        rg = float(np.mean(rg))
        print("relMSE plain %.5g guided %.5g ratio %.3f; losses %s" % (rp, rg, rg / rp, losses[0]))
        assert all(l[-1] < l[0] for l in losses), losses
        assert rg < rp, (rg, rp)
    finally:
        ds.release()
