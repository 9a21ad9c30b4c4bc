"""Pins of the learned-proposal oracle (oracle/guide_ref.py) by mathematics alone (no GPU)."""
import numpy as np
import pytest

from oracle import guide_ref as G


def _params(rng, scale=0.15):
    p = rng.normal(0.0, scale, G.PARAM_COUNT) * G.pad_mask()
    return p


def _batch(rng, n):
    x = rng.uniform(0, 1, (n, 9)).astype(np.float32)
    w = rng.normal(size=(n, 3))
    w /= np.linalg.norm(w, axis=1, keepdims=True)
    wt = rng.exponential(1.0, n)
    return x, w, wt


def test_param_layout():
    assert G.PARAM_COUNT == 304 * 64 + 64 + 3 * (64 * 64 + 64) + 64 * 48 + 48 == 35120
    p = np.arange(G.PARAM_COUNT, dtype=np.float64)
    assert np.array_equal(G.pack(G.unpack(p)), p)


def test_bf16_rounding_known_values():
    vals = np.array([1.0, 1.00390625, 1.005859375, 1.0078125, -2.5, 3.0e38, 1e-40], dtype=np.float32)
    r = G.bf16(vals)
    # 1 + 2^-8 is a tie between 1 and 1 + 2^-7: even mantissa -> 1; 1 + 1.5 * 2^-8 rounds up
    assert r[0] == 1.0 and r[1] == 1.0 and r[2] == 1.0078125 and r[3] == 1.0078125 and r[4] == -2.5
    assert np.isfinite(r[5]) and abs(r[5] - 3.0e38) / 3.0e38 < 2 ** -8
    assert r[6] >= 0.0


def test_encoding_closed_form():
    x = np.array([[0.0, 0.25, 0.5, 0.75, 1.0, 0.125, 0.3, 0.9, 0.6]], dtype=np.float32)
    e = G.encode(x)
    assert e.shape == (1, 304) and np.all(e[:, 297:] == 0)
    assert np.array_equal(e[0, :9], x[0].astype(np.float64))
    # input 1 (0.25), frequency 1: sin(pi/2) = 1, cos = 0; frequency 2 (k = 2): sin(pi) = 0, cos = -1
    assert np.isclose(e[0, 9 + 32 + 0], 1.0) and np.isclose(e[0, 9 + 32 + 1], 0.0, atol=1e-12)
    assert np.isclose(e[0, 9 + 32 + 4], 0.0, atol=1e-12) and np.isclose(e[0, 9 + 32 + 5], -1.0)
    # k = 1 is 2^(1/2) cycles: 0.25 * sqrt(2) of a revolution
    assert np.isclose(e[0, 9 + 32 + 2], np.sin(2 * np.pi * 0.25 * np.sqrt(2)), atol=1e-6)
    assert [float(G.freq(k)) for k in (0, 2, 4, 15)] == [1.0, 2.0, 4.0, float(np.float32(np.sqrt(2)) * 128)]
    # sin^2 + cos^2 = 1 for every pair
    s, c = e[0, 9:297:2], e[0, 10:297:2]
    assert np.allclose(s * s + c * c, 1.0)


def test_backward_matches_finite_differences():
    rng = np.random.default_rng(5)
    p = _params(rng)
    x, w, wt = _batch(rng, 24)
    loss, g = G.gradients(p, x, w, wt, rounding=False)
    assert np.isclose(loss, G.loss_only(p, x, w, wt))
    idx = np.concatenate([rng.choice(np.flatnonzero(G.pad_mask()), 60, replace=False),
                          [G.PARAM_COUNT - 48 - 1, 304 * 64 + 3]])     # a W5 entry, a b1 entry
    eps = 1e-6
    for i in idx:
        dp = np.zeros_like(p)
        dp[i] = eps
        fd = (G.loss_only(p + dp, x, w, wt) - G.loss_only(p - dp, x, w, wt)) / (2 * eps)
        assert abs(fd - g[i]) <= 1e-6 + 1e-4 * abs(fd), (i, fd, g[i])
    assert np.all(g[G.pad_mask() == 0] == 0)


def test_dout_matches_finite_differences():
    rng = np.random.default_rng(6)
    o = rng.normal(0, 1.0, (5, 48))
    o[:, 4::5][:, :8] = rng.uniform(-2.5, 6.5, (5, 8))       # log-kappa inside the clamp
    _, w, wt = _batch(rng, 5)
    _, d = G.loss_and_dout(o, w, wt)
    eps = 1e-6
    for i in range(5):
        for j in range(40):
            op, om = o.copy(), o.copy()
            op[i, j] += eps
            om[i, j] -= eps
            fd = (G.loss_and_dout(op, w, wt)[0] - G.loss_and_dout(om, w, wt)[0]) / (2 * eps)
            assert abs(fd - d[i, j]) <= 1e-7 + 1e-5 * abs(fd)


def _sphere_quadrature(n_theta=400, n_phi=800):
    th = (np.arange(n_theta) + 0.5) * np.pi / n_theta
    ph = (np.arange(n_phi) + 0.5) * 2 * np.pi / n_phi
    T, P = np.meshgrid(th, ph, indexing="ij")
    dirs = np.stack([np.sin(T) * np.cos(P), np.sin(T) * np.sin(P), np.cos(T)], axis=-1).reshape(-1, 3)
    dw = (np.sin(T) * (np.pi / n_theta) * (2 * np.pi / n_phi)).reshape(-1)
    return dirs, dw


@pytest.mark.parametrize("kappa", [0.05, 1.0, 20.0, 200.0])
def test_vmf_normalisation_and_mean(kappa):
    dirs, dw = _sphere_quadrature()
    mu = np.array([0.3, -0.5, 0.81])
    mu /= np.linalg.norm(mu)
    lv = G.log_vmf(np.broadcast_to(mu, (dirs.shape[0], 1, 3)), np.full((dirs.shape[0], 1), kappa), dirs)[:, 0]
    pdf = np.exp(lv)
    assert abs((pdf * dw).sum() - 1.0) < 2e-3
    mean = (pdf[:, None] * dirs * dw[:, None]).sum(axis=0)
    closed = 1.0 / np.tanh(kappa) - 1.0 / kappa                     # mean resultant length
    assert abs(np.linalg.norm(mean) - closed) < 3e-3
    assert np.allclose(mean / np.linalg.norm(mean), mu, atol=1e-3)


def test_mixture_pdf_normalised():
    rng = np.random.default_rng(8)
    o = rng.normal(0, 1.0, (3, 48))
    o[:, 4:40:5] = rng.uniform(-1.0, 3.0, (3, 8))
    dirs, dw = _sphere_quadrature(300, 600)
    for i in range(3):
        q = G.mixture_pdf(np.repeat(o[i:i + 1], dirs.shape[0], axis=0), dirs)
        assert abs((q * dw).sum() - 1.0) < 2e-3


def test_vmf_sampling_matches_pdf():
    rng = np.random.default_rng(9)
    kappa = 6.0
    mu = np.array([0.0, 0.6, 0.8])
    n = 200000
    s = G.sample_vmf(np.broadcast_to(mu, (n, 3)), np.full(n, kappa), rng.uniform(size=n), rng.uniform(size=n))
    assert np.allclose(np.linalg.norm(s, axis=1), 1.0)
    c = s @ mu
    # the density of c = mu.w is kappa e^{kappa (c-1)} / (1 - e^{-2 kappa}) on [-1, 1]
    hist, edges = np.histogram(c, bins=40, range=(-1, 1), density=True)
    mid = 0.5 * (edges[1:] + edges[:-1])
    ref = kappa * np.exp(kappa * (mid - 1)) / (-np.expm1(-2 * kappa))
    big = ref > 0.05
    assert np.all(np.abs(hist[big] - ref[big]) / ref[big] < 0.08)
    # azimuth uniform around mu
    t, b = G.frame(mu[None, :])
    phi = np.arctan2(s @ b[0], s @ t[0])
    h2, _ = np.histogram(phi, bins=16, range=(-np.pi, np.pi))
    assert np.all(np.abs(h2 / (n / 16) - 1) < 0.03)


def test_adam_step_closed_form():
    p = np.array([1.0, -2.0])
    g = np.array([0.5, -0.25])
    p1, m, v = G.adam(p, g, np.zeros(2), np.zeros(2), 1, lr=0.1)
    # first step with bias correction moves each parameter by lr * sign(g) (up to eps)
    assert np.allclose(p1, p - 0.1 * np.sign(g), atol=1e-6)
