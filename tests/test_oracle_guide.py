"""The CPU oracle of the learned proposal, pinned without the GPU:
- each vMF lobe integrates to 1 (adaptive quadrature), and so does a mixture
  of decoded lobes (quadrature over the sphere);
- the lobe sampler is the inverse of the numerically integrated CDF;
- the analytic FP64 gradient equals central finite differences of the loss
  for coordinates in every layer (catches a dropped term, a wrong index or a
  missing ReLU mask);
- bf16 rounding is round-to-nearest-even on 16 bits;
- the input encoding is sin/cos(2^k pi x)."""
import numpy as np
import pytest
from scipy import integrate

import oracle


def rand_params(rng, scale=1.0):
    P = np.zeros(oracle.GUIDE_PARAMS)
    for k in range(5):
        fan_in = 53 if k == 0 else 64
        rows = 64 if k < 4 else 40
        W = rng.uniform(-1, 1, (rows, fan_in)) * np.sqrt(6.0 / fan_in) * (scale if k < 4 else 0.5)
        base = k * 4160
        Wf = np.zeros((64, 64))
        Wf[:rows, :fan_in] = W
        P[base:base + 4096] = Wf.ravel()
        P[base + 4096:base + 4096 + rows] = rng.uniform(-0.2, 0.2, rows)
    return P


def rand_records(rng, n):
    rec = np.zeros((n, 68), np.float32)
    for i in range(n):
        nrm = rng.normal(size=3)
        nrm /= np.linalg.norm(nrm)
        rec[i, :64] = oracle.guide_encode(rng.uniform(0, 1, 3), nrm, rng.uniform(), rng.uniform(0, 0.5))
        d = rng.normal(size=3)
        rec[i, 64:67] = d / np.linalg.norm(d)
        rec[i, 67] = rng.uniform(0.2, 2.0)
    return rec


@pytest.mark.parametrize("kappa", [0.05, 0.5, 1.0, 10.0, 100.0, 1096.6])
def test_vmf_lobe_integrates_to_one(kappa):
    C = oracle.vmf_norm(kappa)
    # integral over the sphere = 2 pi int_{-1}^{1} C exp(kappa (t - 1)) dt, in s = 1 - t
    val, err = integrate.quad(lambda s: C * np.exp(-kappa * s), 0.0, 2.0, points=[min(1.0, 10.0 / kappa)],
                              epsabs=1e-13, epsrel=1e-11, limit=200)
    assert abs(2 * np.pi * val - 1.0) < 1e-8


def test_mixture_pdf_integrates_to_one():
    rng = np.random.default_rng(3)
    raw = rng.normal(size=40)
    for l in range(8):
        raw[5 * l + 3] = rng.uniform(-3, 2.5)   # concentrations up to e^2.5 ~ 12
    nt, nphi = 1600, 1600
    t = -1 + (np.arange(nt) + 0.5) * 2 / nt
    phi = (np.arange(nphi) + 0.5) * 2 * np.pi / nphi
    T, PH = np.meshgrid(t, phi, indexing="ij")
    st = np.sqrt(1 - T * T)
    dirs = np.stack([st * np.cos(PH), st * np.sin(PH), T], -1).reshape(-1, 3)
    # vectorised restatement would retype the formula; sample the oracle on a coarse subset instead
    idx = rng.choice(dirs.shape[0], 4000, replace=False)
    p_or = oracle.vmf_mixture_pdf(raw, dirs[idx])
    # decoded lobes for the full quadrature (normalised means, softmax weights)
    mu, kap, w = [], [], []
    for l in range(8):
        m = raw[5 * l:5 * l + 3]
        mu.append(m / np.linalg.norm(m))
        kap.append(np.exp(np.clip(raw[5 * l + 3], -3, 7)))
        w.append(np.exp(raw[5 * l + 4]))
    w = np.array(w) / np.sum(w)
    p = np.zeros(dirs.shape[0])
    for l in range(8):
        C = oracle.vmf_norm(kap[l])
        p += w[l] * C * np.exp(kap[l] * (dirs @ mu[l] - 1.0))
    assert np.allclose(p[idx], p_or, rtol=1e-9)
    total = p.sum() * (2.0 / nt) * (2 * np.pi / nphi)
    assert abs(total - 1.0) < 1e-4


@pytest.mark.parametrize("kappa", [0.02, 0.7, 5.0, 80.0])
def test_vmf_sampler_inverts_the_cdf(kappa):
    C = oracle.vmf_norm(kappa)
    for u in np.linspace(0.001, 0.999, 23):
        t = oracle.vmf_sample_cos(kappa, u)
        assert -1.0 <= t <= 1.0
        F, _ = integrate.quad(lambda tt: 2 * np.pi * C * np.exp(kappa * (tt - 1.0)), -1.0, t, epsabs=1e-12,
                              epsrel=1e-10, limit=200)
        assert abs(F - u) < 2e-6, (kappa, u, t, F)


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(11)
    P = rand_params(rng)
    rec = rand_records(rng, 6)
    loss, g = oracle.guide_gradient(P, rec, bf16=False)
    assert np.isfinite(loss) and loss > 0
    coords = []
    for k in range(5):
        base = k * 4160
        rows = 64 if k < 4 else 40
        cols = 53 if k == 0 else 64
        for _ in range(10):
            coords.append(base + rng.integers(rows) * 64 + rng.integers(cols))
        for _ in range(3):
            coords.append(base + 4096 + rng.integers(rows))
    worst = 0.0
    for c in coords:
        h = 1e-6 * max(1.0, abs(P[c]))
        Pp, Pm = P.copy(), P.copy()
        Pp[c] += h
        Pm[c] -= h
        lp, _ = oracle.guide_gradient(Pp, rec, bf16=False, want_grad=False)
        lm, _ = oracle.guide_gradient(Pm, rec, bf16=False, want_grad=False)
        fd = (lp - lm) / (2 * h)
        err = abs(fd - g[c]) / max(1e-4, abs(fd), abs(g[c]))
        worst = max(worst, err)
        assert err < 1e-5, (c, fd, g[c])
    # the padded parts carry no gradient
    assert np.all(g[0 * 4160:4096].reshape(64, 64)[:, 53:] == 0)
    assert np.all(g[4 * 4160:4 * 4160 + 4096].reshape(64, 64)[40:, :] == 0)
    assert np.all(g[4 * 4160 + 4096 + 40:5 * 4160] == 0)


def test_bf16_rounding_is_nearest_even():
    assert oracle.bf16_round(1.0) == 1.0
    assert oracle.bf16_round(1.0 + 2 ** -8) == 1.0                 # halfway, even below
    assert oracle.bf16_round(1.0 + 3 * 2 ** -8) == 1.0 + 2 ** -6   # halfway, even above
    assert oracle.bf16_round(1.0 + 2 ** -8 + 2 ** -12) == 1.0 + 2 ** -7
    assert oracle.bf16_round(-3.0e38) == -3.0e38 or abs(oracle.bf16_round(-3.0e38)) > 2.9e38


def test_encoding_is_sin_cos_bands():
    pos = np.array([0.1, 0.55, 0.93], np.float32).astype(np.float64)   # the encoder takes fp32 positions
    x = oracle.guide_encode(pos, [0, 0, 1], 0.25, 0.3)
    for a in range(3):
        for k in range(8):
            assert abs(x[16 * a + 2 * k] - np.sin(np.pi * 2 ** k * pos[a])) < 1e-6
            assert abs(x[16 * a + 2 * k + 1] - np.cos(np.pi * 2 ** k * pos[a])) < 1e-6
    assert list(x[48:53]) == [0.0, 0.0, 1.0, 0.25, pytest.approx(0.3)]
    assert np.all(x[53:] == 0)


def test_bf16_emulation_changes_little():
    rng = np.random.default_rng(5)
    P = rand_params(rng)
    rec = rand_records(rng, 16)
    a = oracle.guide_forward(P, rec[:, :64], bf16=False)
    b = oracle.guide_forward(P, rec[:, :64], bf16=True)
    rel = np.linalg.norm(a - b) / np.linalg.norm(a)
    assert 1e-5 < rel < 5e-2


def numpy_forward(P, x):
    """The network written with numpy matrix products: four hidden layers of
    64 with ReLU, then the 40-output head; each of the 5 parameter blocks is
    a row-major 64 x 64 weight (out, in) followed by 64 biases."""
    h = x.astype(np.float64)
    for k in range(5):
        blk = P[k * 4160:(k + 1) * 4160]
        W, b = blk[:4096].reshape(64, 64), blk[4096:]
        z = h @ W.T + b
        h = np.maximum(z, 0.0) if k < 4 else z
    return h[:, :40]


def numpy_mixture_pdf(raw, d):
    """vMF mixture density of one head output: lobe i is raw[5i:5i+5] =
    (mu direction unnormalised, log kappa clamped to [-3, 7], logit of the
    weight); weights are the softmax of the logits; a lobe's density is
    kappa / (2 pi (1 - e^{-2 kappa})) e^{kappa (mu.d - 1)}."""
    r = raw.reshape(8, 5)
    mu = r[:, :3] / np.linalg.norm(r[:, :3], axis=1, keepdims=True)
    kappa = np.exp(np.clip(r[:, 3], -3.0, 7.0))
    w = np.exp(r[:, 4] - r[:, 4].max())
    w /= w.sum()
    lobe = kappa / (2 * np.pi * (1 - np.exp(-2 * kappa))) * np.exp(kappa * (mu @ d - 1.0))
    return float(np.dot(w, lobe))


def test_forward_matches_numpy_matmul():
    """oracle.guide_forward (FP64, no bf16) against numpy matrix products at
    random parameters and inputs; a transposed weight or a missing bias
    would move the outputs by O(1)."""
    rng = np.random.default_rng(5)
    P = rng.normal(0.0, 0.25, oracle.GUIDE_PARAMS)
    x = rng.uniform(-1.0, 1.0, (64, 64)).astype(np.float32)
    got = oracle.guide_forward(P, x, bf16=False)
    want = numpy_forward(P, x)
    assert np.allclose(got, want, rtol=1e-12, atol=1e-12 * np.abs(want).max())
    # the comparison has power: a transposed layer differs
    Pt = P.copy()
    Pt[4160:4160 + 4096] = P[4160:4160 + 4096].reshape(64, 64).T.ravel()
    assert np.abs(numpy_forward(Pt, x) - want).max() > 1e-3 * np.abs(want).max()


def test_decoded_mixture_matches_numpy():
    rng = np.random.default_rng(6)
    for _ in range(20):
        raw = rng.normal(0.0, 1.5, 40)
        dirs = rng.normal(size=(8, 3))
        dirs /= np.linalg.norm(dirs, axis=1, keepdims=True)
        got = oracle.vmf_mixture_pdf(raw, dirs)
        want = np.array([numpy_mixture_pdf(raw, d) for d in dirs])
        assert np.allclose(got, want, rtol=1e-12)
