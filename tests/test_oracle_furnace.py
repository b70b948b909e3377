"""The oracle's multi-bounce path loop pinned by closed forms (tests/furnace.py):
the Lambert furnace at depth limits 0, 1, 3 and 8 with Russian roulette off
and on, the lossless-glass furnace, and the zero-variance property of
etaScale inside lossless glass. Each check is also shown to reject a planted
mistake of the kind it is there to catch (oracle OR_MUT_*): roulette
survivors divided by q, emission counted at every hit, the depth break one
bounce late, etaScale never updated, emission after specular bounces
dropped. PathIntegrator::Li: integrators/pathintegrator.cpp:12-87."""
import numpy as np
import pytest

import oracle
from furnace import (LE, box_triangles, expected_glass_parity, expected_parity, furnace, glass_pixels, image_mean_z,
                     parity_xyz, spd_bins)

NO_RR = 1 << 20
Z_OK = 4.5        # |z| bound of a correct estimator (per channel; p ~ 7e-6 each)
Z_CAUGHT = 8.0    # |z| a planted mistake must exceed


def test_box_winding_and_closure():
    t = box_triangles((-1, -1, -1), (1, 1, 1), inward=True).astype(np.float64)
    n = np.cross(t[:, 1] - t[:, 0], t[:, 2] - t[:, 0])
    c = t.mean(1)
    assert np.all(np.einsum("ij,ij->i", n, -c) > 0), "inward normals point to the centre"
    # closed: every edge is shared by exactly two triangles, traversed in opposite directions
    edges = {}
    for tri in t:
        for a, b in ((0, 1), (1, 2), (2, 0)):
            key = (tuple(tri[a]), tuple(tri[b]))
            edges[key] = edges.get(key, 0) + 1
    for (a, b), k in edges.items():
        assert k == 1 and edges.get((b, a), 0) == 1
    assert np.isclose(np.linalg.norm(n, axis=1).sum() / 2, 24.0)   # surface area of [-1, 1]^3
    g = box_triangles((-0.5, -0.5, -0.7), (0.5, 0.5, -0.1), inward=False).astype(np.float64)
    ng = np.cross(g[:, 1] - g[:, 0], g[:, 2] - g[:, 0])
    assert np.all(np.einsum("ij,ij->i", ng, g.mean(1) - np.array([0, 0, -0.4])) > 0)


def test_constant_spectra_are_constant():
    """The constant SPDs average to their value in every bin (to float
    rounding), so the closed forms' per-bin series are those of rho and 1."""
    for v in (0.0, 0.5, 0.8, 1.0):
        b = spd_bins(v)
        assert np.all(np.abs(b - v) <= 2e-7 * max(v, 1.0)), v


def render(gen, spp, D, rr_start=NO_RR, seed=11, mutation=oracle.MUT_NONE, depths=None, rr_threshold=1.0):
    s, q, st = oracle.Scene(gen).render(spp, D, seed=seed, threads=8, rr_start=rr_start, depths=depths,
                                        rr_threshold=rr_threshold, mutation=mutation)
    return s, q


# roulette from bounce 1 on (rr_start 0) and after bounce 3 (the reference's
# rule); at D = 1 the break comes before any roulette, so D = 2 is the
# smallest depth where it acts
CASES = [(0, NO_RR), (1, NO_RR), (3, NO_RR), (8, NO_RR), (2, 0), (3, 0), (8, 0), (8, 3)]


@pytest.mark.parametrize("D,rr_start", CASES)
def test_lambert_furnace_closed_form(D, rr_start):
    rho, spp = 0.8, (256 if D >= 8 else 128)
    gen = furnace(rho)
    s, q = render(gen, spp, D, rr_start)
    exp = expected_parity(rho, D)
    z, mean, sigma = image_mean_z(s, q, spp, exp)
    assert np.all(np.abs(z) < Z_OK), (z, mean, exp)
    # the closed form of the neighbouring depths is rejected: the check can
    # tell depth D from D - 1 and D + 1 at this sample count
    for Dn in (D - 1, D + 1):
        if Dn < 0:
            continue
        zn = (mean - expected_parity(rho, Dn)) / sigma
        assert np.all(np.abs(zn) > Z_CAUGHT), (Dn, zn)


@pytest.mark.parametrize("mutation,D,rr_start", [(oracle.MUT_RR_DIV_Q, 8, 0), (oracle.MUT_EMIT_ALWAYS, 3, NO_RR),
                                                 (oracle.MUT_DEPTH_LATE, 3, NO_RR), (oracle.MUT_DEPTH_LATE, 8, 3)])
def test_lambert_furnace_rejects_planted_mistakes(mutation, D, rr_start):
    rho, spp = 0.8, 128
    s, q = render(furnace(rho), spp, D, rr_start, mutation=mutation)
    z, _, _ = image_mean_z(s, q, spp, expected_parity(rho, D))
    assert np.all(np.abs(z) > Z_CAUGHT), z


GLASS_DEPTHS = (64, 64, 64)


def test_glass_furnace_closed_form():
    rho, spp = 0.5, 128
    gen = furnace(rho, glass=True)
    inside, outside = glass_pixels(gen)
    assert inside.sum() >= 30 and outside.sum() >= 100
    s, q = render(gen, spp, 0, rr_start=3, depths=GLASS_DEPTHS)
    exp = expected_glass_parity(rho)
    # the depth limit leaves out at most rho^(diffuse vertices) of the series:
    # negligible at 64 vertices
    for mask in (inside, outside, None):
        z, mean, _ = image_mean_z(s, q, spp, exp, mask)
        assert np.all(np.abs(z) < Z_OK), (z, mean, exp)


def test_glass_furnace_rejects_dropped_specular_emission():
    rho, spp = 0.5, 128
    gen = furnace(rho, glass=True)
    inside, _ = glass_pixels(gen)
    s, q = render(gen, spp, 0, rr_start=3, depths=GLASS_DEPTHS, mutation=oracle.MUT_NO_SPECULAR_EMIT)
    z, _, _ = image_mean_z(s, q, spp, expected_glass_parity(rho), inside)
    assert np.all(np.abs(z) > Z_CAUGHT), z


def glass_black_walls(mutation=oracle.MUT_NONE, spp=64):
    gen = furnace(0.0, glass=True)
    s, q = render(gen, spp, 0, rr_start=0, depths=GLASS_DEPTHS, rr_threshold=0.99, mutation=mutation)
    inside, outside = glass_pixels(gen)
    m = s.astype(np.float64) / spp
    var = np.maximum(q.astype(np.float64) / spp - m * m, 0.0)
    return m, var, inside, outside


def test_eta_scale_keeps_lossless_glass_noise_free():
    """Every sample through lossless glass returns exactly Le (the walls are
    black emitters): the oracle's 60-bin spectra make each sample
    deterministic, so the glass pixels' sample variance is zero to float
    rounding, and their mean is Le's XYZ."""
    m, var, inside, outside = glass_black_walls()
    want = parity_xyz(spd_bins(1.0) * LE)
    for mask in (inside, outside):
        assert np.allclose(m[mask], want, rtol=2e-5), (m[mask].min(0), m[mask].max(0), want)
        # the float32 film's sum of squares leaves sqrt(eps) ~ 3.5e-4 of
        # cancellation in q/n - m^2; roulette inside the glass would give O(1)
        assert np.all(np.sqrt(var[mask]) <= 2e-3 * want), np.sqrt(var[mask]).max(0)


def test_eta_scale_check_rejects_missing_eta_scale():
    m, var, inside, _ = glass_black_walls(oracle.MUT_NO_ETASCALE)
    want = parity_xyz(spd_bins(1.0) * LE)
    rel_sd = np.sqrt(var[inside]) / want
    assert np.median(rel_sd[:, 1]) > 0.1, np.median(rel_sd[:, 1])
