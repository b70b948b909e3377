"""The render path's sampler (the role of Sampler::get_1d/get_2d,
sampling/sampler.h:14-18; the reference's Whitenoise draws from mt19937,
samplers/whitenoise.h:11-29) on the device through fl_sampler_draw:
uniformity per dimension (chi-square over 256 bins), independence of adjacent
dimensions, of the same dimension at consecutive bounces, samples and pixels
(chi-square over 32 x 32 cells), moments and correlations, the 24-bit grid and
[0, 1) range, determinism and seed dependence. The statistics are shown to
reject sequences with a planted flaw (a dimension that copies a shifted
neighbour, a slight non-uniformity)."""
import numpy as np
import pytest
from scipy import stats

pytestmark = pytest.mark.gpu

P_MIN = 1e-6          # smallest accepted p-value of each test
N_DIMS = 12           # dimensions the shading kernel draws per vertex (0..11)
CAMERA_BOUNCE = 0xFF  # the key of camera-ray generation


def keys(pixels, samples, bounce):
    p, s = np.meshgrid(np.asarray(pixels, np.uint32), np.asarray(samples, np.uint32), indexing="ij")
    return np.stack([p.ravel(), s.ravel(), np.full(p.size, bounce, np.uint32)], 1)


def chi2_uniform_p(u, bins=256):
    h = np.bincount(np.minimum((u * bins).astype(np.int64), bins - 1), minlength=bins)
    return stats.chisquare(h).pvalue


def chi2_pair_p(a, b, bins=32):
    ia = np.minimum((a * bins).astype(np.int64), bins - 1)
    ib = np.minimum((b * bins).astype(np.int64), bins - 1)
    h = np.bincount(ia * bins + ib, minlength=bins * bins)
    return stats.chisquare(h).pvalue


def check_sequence(U, where=""):
    """U: (n, d) draws of n keys. Raises AssertionError on a failed test."""
    n = U.shape[0]
    for d in range(U.shape[1]):
        u = U[:, d].astype(np.float64)
        assert chi2_uniform_p(u) > P_MIN, ("uniformity", where, d)
        assert abs(u.mean() - 0.5) < 5 * np.sqrt(1 / 12 / n), ("mean", where, d)
        assert abs(u.var() - 1 / 12) < 5 * np.sqrt(1 / 180 / n), ("variance", where, d)
    for d in range(U.shape[1] - 1):
        a, b = U[:, d].astype(np.float64), U[:, d + 1].astype(np.float64)
        assert chi2_pair_p(a, b) > P_MIN, ("adjacent dimensions", where, d)
        assert abs(np.corrcoef(a, b)[0, 1]) < 5 / np.sqrt(n), ("correlation", where, d)


@pytest.fixture(scope="module")
def draws(ctx):
    pixels, samples = np.arange(1024), np.arange(1024)
    out = {}
    for b in (0, 1, 2, CAMERA_BOUNCE):
        out[b] = ctx.sampler_draw(7, keys(pixels, samples, b), N_DIMS).reshape(1024, 1024, N_DIMS)
    return out


def test_range_and_grid(draws):
    for U in draws.values():
        assert U.min() >= 0.0 and U.max() < 1.0
        assert np.all((U * 2.0 ** 24) == np.floor(U * 2.0 ** 24))


@pytest.mark.parametrize("bounce", [0, 1, CAMERA_BOUNCE])
def test_uniform_and_independent_dimensions(draws, bounce):
    check_sequence(draws[bounce].reshape(-1, N_DIMS), "bounce %d" % bounce)


def test_independent_across_bounces_samples_and_pixels(draws):
    U0, U1, U2 = draws[0], draws[1], draws[2]
    for d in range(N_DIMS):
        assert chi2_pair_p(U0[..., d].ravel(), U1[..., d].ravel()) > P_MIN, ("bounces 0/1", d)
        assert chi2_pair_p(U1[..., d].ravel(), U2[..., d].ravel()) > P_MIN, ("bounces 1/2", d)
        assert chi2_pair_p(U0[:, :-1, d].ravel(), U0[:, 1:, d].ravel()) > P_MIN, ("samples", d)
        assert chi2_pair_p(U0[:-1, :, d].ravel(), U0[1:, :, d].ravel()) > P_MIN, ("pixels", d)
        assert chi2_pair_p(draws[CAMERA_BOUNCE][..., d].ravel(), U0[..., d].ravel()) > P_MIN, ("camera/bounce 0", d)


def test_deterministic_and_seeded(ctx):
    k = keys(np.arange(64), np.arange(64), 3)
    a, b = ctx.sampler_draw(11, k, 4), ctx.sampler_draw(11, k, 4)
    assert np.array_equal(a, b)
    c = ctx.sampler_draw(12, k, 4)
    assert np.mean(a == c) < 0.01
    hi = ctx.sampler_draw(11 + (1 << 32), k, 4)   # the seed's upper half enters too
    assert np.mean(a == hi) < 0.01


def test_statistics_reject_planted_flaws(draws):
    U = draws[0].reshape(-1, N_DIMS).astype(np.float64).copy()
    bad = U.copy()
    bad[:, 5] = np.mod(bad[:, 4] + 0.37, 1.0)            # a dimension that copies its neighbour
    with pytest.raises(AssertionError):
        check_sequence(bad)
    skew = U.copy()
    skew[:, 2] = skew[:, 2] ** 1.01                       # 1% non-uniformity
    with pytest.raises(AssertionError):
        check_sequence(skew)
