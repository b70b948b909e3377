"""The guiding network's matrix-core kernels against the FP64 oracle, which
rounds to bf16 where the GPU does (so only fp32 accumulation order differs),
plus training progress on a known target and divergence handling."""
import numpy as np
import pytest

import flgpu
import oracle

pytestmark = pytest.mark.gpu


def encoded_records(rng, n, target=False):
    rec = np.zeros((n, 68), np.float32)
    for i in range(n):
        pos = rng.uniform(0, 1, 3)
        nrm = rng.normal(size=3)
        nrm /= np.linalg.norm(nrm)
        rec[i, :64] = oracle.guide_encode(pos, nrm, rng.uniform(), rng.uniform(0, 0.5))
        if target:
            # directions from a vMF (kappa 20) whose axis depends on the position
            mu = np.array([np.sin(2 * np.pi * pos[0]), np.cos(2 * np.pi * pos[1]), 0.6])
            mu /= np.linalg.norm(mu)
            t = oracle.vmf_sample_cos(20.0, rng.uniform())
            phi = rng.uniform(0, 2 * np.pi)
            a = np.cross(mu, [1, 0, 0] if abs(mu[0]) < 0.9 else [0, 1, 0])
            a /= np.linalg.norm(a)
            b = np.cross(mu, a)
            s = np.sqrt(max(0.0, 1 - t * t))
            d = mu * t + a * s * np.cos(phi) + b * s * np.sin(phi)
            rec[i, 67] = 1.0
        else:
            d = rng.normal(size=3)
            rec[i, 67] = rng.uniform(0.2, 2.0)
        rec[i, 64:67] = d / np.linalg.norm(d)
    return rec


def test_forward_matches_oracle(ctx):
    rng = np.random.default_rng(1)
    with ctx.guide(seed=3) as g:
        P = g.params()[0].astype(np.float64)
        x = encoded_records(rng, 300)[:, :64]   # not a multiple of 64: exercises the tail block
        got = g.forward(x)
        ref = oracle.guide_forward(P, x, bf16=True)
    rel = np.linalg.norm(got - ref) / np.linalg.norm(ref)
    print("forward rel L2 %.2e" % rel)
    assert rel < 1e-4


def test_gradient_matches_oracle(ctx):
    rng = np.random.default_rng(2)
    with ctx.guide(seed=4) as g:
        P = g.params()[0].astype(np.float64)
        rec = encoded_records(rng, 300)
        loss, grad = g.gradient(rec)
    loss_o, grad_o = oracle.guide_gradient(P, rec, bf16=True)
    assert abs(loss - loss_o) / abs(loss_o) < 1e-4
    for k in range(5):
        sl = slice(k * 4160, (k + 1) * 4160)
        rel = np.linalg.norm(grad[sl] - grad_o[sl]) / np.linalg.norm(grad_o[sl])
        print("matrix %d: gradient rel L2 %.2e" % (k, rel))
        assert rel < 1e-4
    # and far from the exact fp64 gradient of a transposed weight (a layout mistake)
    Pt = P.copy()
    W1 = Pt[4160:4160 + 4096].reshape(64, 64)
    Pt[4160:4160 + 4096] = W1.T.ravel()
    _, grad_t = oracle.guide_gradient(Pt, rec, bf16=True)
    assert np.linalg.norm(grad - grad_t) / np.linalg.norm(grad_t) > 0.1


def test_training_learns_a_known_distribution(ctx):
    rng = np.random.default_rng(3)
    train = encoded_records(rng, 8192, target=True)
    held = encoded_records(rng, 2048, target=True)
    with ctx.guide(seed=5, lr=3e-3) as g:
        l0, _ = g.gradient(held)
        for step in range(300):
            idx = rng.integers(0, train.shape[0], 2048)
            rc, _ = g.train_step(train[idx])
            assert rc == flgpu.FL_OK
        l1, _ = g.gradient(held)
        steps, last, div = g.stats()
    nll0, nll1 = l0 / held.shape[0], l1 / held.shape[0]
    print("held-out NLL %.3f -> %.3f (uniform sphere: %.3f)" % (nll0, nll1, np.log(4 * np.pi)))
    assert steps == 300 and div == 0
    assert nll1 < nll0 - 1.0
    assert nll1 < np.log(4 * np.pi) - 1.0   # clearly better than a uniform proposal


def test_divergence_is_detected_and_cleared(ctx):
    rng = np.random.default_rng(4)
    rec = encoded_records(rng, 256)
    with ctx.guide(seed=6) as g:
        P = g.params()[0].copy()
        bad = P.copy()
        bad[4160 + 7] = np.nan
        g.set_params(bad)
        rc, loss = g.train_step(rec)
        assert rc == flgpu.FL_E_NUMERIC
        assert g.stats()[2] == 1
        assert np.isnan(g.params()[0][4160 + 7])   # the update was skipped
        g.set_params(P)
        assert g.stats()[2] == 0
        rc, loss = g.train_step(rec)
        assert rc == flgpu.FL_OK and np.isfinite(loss)
