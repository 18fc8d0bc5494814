"""Pins of the oracle's restatement of the learned proposal (the HIP path's
guiding network and its vMF mixture), from the mathematics alone -- the
reference renderer has no guiding:

  * the vMF density integrates to 1 over the sphere, for concentrations from
    nearly uniform (0.05) to sharp (1000);
  * vMF samples follow the density: the cosine to the mean direction has the
    closed-form CDF (e^{k(c-1)} - e^{-2k}) / (1 - e^{-2k}) (KS test), and the
    azimuth is uniform;
  * the 8-lobe mixture of random network outputs integrates to 1;
  * the analytic gradient of the training loss (backpropagation through the
    mixture and the three layers) equals central differences of the loss.
"""
import math

import numpy as np
import pytest

import oracle_lib as ol

BOX = np.array([0, 0, -555, 555, 555, 0], np.float32)


def fib_sphere(n):
    """n nearly uniform directions (Fibonacci lattice)."""
    i = np.arange(n) + 0.5
    z = 1.0 - 2.0 * i / n
    r = np.sqrt(1.0 - z * z)
    ph = math.pi * (3.0 - math.sqrt(5.0)) * i
    return np.stack([r * np.cos(ph), r * np.sin(ph), z], 1)


@pytest.mark.parametrize("kappa", [0.05, 1.0, 10.0, 100.0, 1000.0])
def test_vmf_pdf_normalised(kappa):
    mu = np.array([0.3, -0.5, 0.81])
    mu /= np.linalg.norm(mu)
    # the density depends on c = mu.w only: 2 pi * integral over c in [-1, 1],
    # Gauss-Legendre nodes mapped onto directions at those cosines
    x, wq = np.polynomial.legendre.leggauss(4000)
    if kappa > 50:  # concentrate the nodes near c = 1
        x = 1.0 - (1.0 - x) * (20.0 / kappa)
        wq = wq * (20.0 / kappa)
    t = np.cross(mu, [1.0, 0.0, 0.0])
    t /= np.linalg.norm(t)
    tot = 0.0
    for c, wt in zip(x, wq):
        w = c * mu + math.sqrt(max(0.0, 1.0 - c * c)) * t
        tot += wt * ol.vmf_pdf(mu, kappa, w)
    assert abs(2 * math.pi * tot - 1.0) < 1e-6, 2 * math.pi * tot


@pytest.mark.parametrize("kappa", [0.5, 8.0, 200.0])
def test_vmf_sampling_matches_density(kappa):
    g = np.random.default_rng(11)
    mu = np.array([-0.2, 0.9, 0.1])
    mu /= np.linalg.norm(mu)
    n = 20000
    u = g.random((n, 2))
    s = np.array([ol.vmf_sample(mu, kappa, a, b) for a, b in u])
    assert np.allclose(np.linalg.norm(s, axis=1), 1.0, atol=1e-12)
    c = np.sort(s @ mu)
    cdf = (np.exp(kappa * (c - 1.0)) - math.exp(-2 * kappa)) / (1.0 - math.exp(-2 * kappa))
    ks = np.max(np.abs(cdf - (np.arange(n) + 0.5) / n))
    assert ks < 1.63 / math.sqrt(n)  # 1% level of the Kolmogorov distribution
    # azimuth about mu is uniform
    t = np.cross(mu, [1.0, 0.0, 0.0])
    t /= np.linalg.norm(t)
    b = np.cross(mu, t)
    ph = np.sort(np.mod(np.arctan2(s @ b, s @ t), 2 * math.pi)) / (2 * math.pi)
    assert np.max(np.abs(ph - (np.arange(n) + 0.5) / n)) < 1.63 / math.sqrt(n)


def test_mixture_normalised():
    g = np.random.default_rng(5)
    dirs = fib_sphere(200000)
    for trial in range(3):
        o = g.normal(0, 1.0, 40)
        o[3::5] = g.uniform(-2, 4, 8)  # log kappa up to e^4 ~ 55
        p = np.array([ol.guide_mix_pdf(o, w) for w in dirs[::4]])
        assert abs(p.mean() * 4 * math.pi - 1.0) < 2e-3


def test_gradient_central_differences():
    """Unquantised (FP64) loss: analytic gradient vs central differences on
    parameters of every layer."""
    g = np.random.default_rng(3)
    params = ol.guide_random_params(7)
    n = 24
    recs = np.zeros((n, 10), np.float32)
    recs[:, 0:3] = g.uniform([0, 0, -555], [555, 555, 0], (n, 3))
    nrm = g.normal(size=(n, 3))
    recs[:, 3:6] = nrm / np.linalg.norm(nrm, axis=1, keepdims=True)
    d = g.normal(size=(n, 3))
    recs[:, 6:9] = d / np.linalg.norm(d, axis=1, keepdims=True)
    recs[:, 9] = g.uniform(0.1, 2.0, n)
    grad, loss = ol.guide_gradient(params, BOX, recs, quantize=False)
    assert np.isfinite(loss) and loss > 0
    # parameters of each block: W1 (used rows), b1, W2, b2, W3 (used columns), b3 (used)
    picks = [g.integers(0, 102) * 64 + g.integers(0, 64) for _ in range(8)]
    picks += [8192 + g.integers(0, 64) for _ in range(4)]
    picks += [8256 + g.integers(0, 4096) for _ in range(8)]
    picks += [12352 + g.integers(0, 64) for _ in range(4)]
    picks += [12416 + g.integers(0, 64) * 48 + g.integers(0, 40) for _ in range(8)]
    picks += [15488 + k for k in range(40)]
    checked = 0
    for p in picks:
        base = params[p]
        h = 1e-4 * max(1.0, abs(float(base)))
        pp, pm = params.copy(), params.copy()
        pp[p] = base + np.float32(h)
        pm[p] = base - np.float32(h)
        dh = float(pp[p]) - float(pm[p])  # the step the float parameters really take
        _, lp = ol.guide_gradient(pp, BOX, recs, quantize=False)
        _, lm = ol.guide_gradient(pm, BOX, recs, quantize=False)
        fd = (lp - lm) / dh
        assert abs(fd - grad[p]) <= 1e-4 * max(abs(fd), abs(grad[p])) + 1e-8, (p, fd, grad[p])
        checked += 1
    assert checked == len(picks)


def test_padding_has_no_gradient():
    params = ol.guide_random_params(9)
    recs = np.array([[100, 200, -300, 0, 1, 0, 0.6, 0.8, 0, 1.0]], np.float32)
    grad, _ = ol.guide_gradient(params, BOX, recs)
    g = grad[12416:12416 + 64 * 48].reshape(64, 48)
    assert np.all(g[:, 40:] == 0) and np.all(grad[15488 + 40:] == 0)
    assert np.all(grad[102 * 64:128 * 64] == 0)  # padded inputs are zero
