"""The oracle's StandardSurfaceMaterial lobe stack (standardsurface.cpp:72-159,
bsdf.cpp:12-151) pinned by closed forms and invariants, not by retyping it:

* kdroughness 0: OrenNayar (sigma 0) is Lambert, f = kd w / pi, pdf cos/pi;
  thin walled, DiffuseTransmission is the same on the far side and 0 near;
* OrenNayar with the view along the normal: f = kd w A / pi with the
  published A = 1 - s^2 / (2 (s^2 + 0.33)) (the B term vanishes there);
* reciprocity f(wo, wi) = f(wi, wo) of every reflection stack;
* thin film along the normal (f = F / (4 pi a^2), GGX D(n) = 1/(pi a^2),
  G = 1): a film of the substrate's index gives the single-interface
  F = ((n - 1)/(n + 1))^2 at every wavelength; a film of index between 1
  and the substrate's gives the Airy reflectance
  (r01^2 + r12^2 + 2 r01 r12 cos phi) / (1 + r01^2 r12^2 + 2 r01 r12 cos phi),
  phi = 4 pi n d / lambda;
* clearcoat along the normal: f = coat w Fr D / 16 (Smith G1 terms 1/2 each,
  Burley's GTR1 D(0) = (a^2 - 1) / (pi ln(a^2) a^2), gloss -> a by
  lerp(gloss, .1, .001));
* with wo along the normal, the integral of pdf over wi is the half-vector
  mass inside 45 degrees: (1/sqrt 2)^(1/r + 2) for Estevez sheen,
  1 - ln((1 + a^2)/2) / ln(a^2) for GTR1;
* E[f cos / pdf] over a stratified sample grid equals the integral of f cos
  over the sphere (sampling, pdf and evaluate agree) for every lobe whose
  sampler draws the density its pdf reports, and for a stack (the one-sample
  lobe mixture of bsdf.cpp:54-126); GGX lobes are asked for the
  visible-normal pdf (FL_RENDER_VNDF_PDF), the density their VNDF sampler
  draws (the reference's default D cos pdf differs away from the normal, as
  the single-lobe conductor tests show); the sheen lobe's draws follow
  cosine-weighted half vectors (its reported pdf is D cos, which it does not
  sample: a reference property kept as is);
* BSDF::sample's pdf and f equal BSDF::pdf and evaluate at the sampled
  direction; with the specular FresnelSpecular lobe in the stack the
  sample's pdf divides by one more component than the non-specular pdf.
"""
import numpy as np
import pytest

import flgpu
import oracle
from stdsurf import LANES, MATCHED, material, queries, random_dirs, sphere_grid, strata, surface, unit_rows

KD = np.array(LANES["kd"])
COAT = np.array(LANES["coat"])
LAM = np.array(LANES["lam"])


def run(m, q):
    return oracle.bsdf_standard(m, q, 64)


def _same_and_opposite(rng, n):
    wo, wi = random_dirs(rng, n), random_dirs(rng, n)
    wo[:, 2] = np.abs(wo[:, 2])
    same = wi.copy()
    same[:, 2] = np.abs(same[:, 2])
    opp = wi.copy()
    opp[:, 2] = -np.abs(opp[:, 2])
    return wo, same, opp


def test_oren_nayar_sigma_zero_is_lambert():
    rng = np.random.default_rng(0)
    wo, same, opp = _same_and_opposite(rng, 400)
    m = surface(kd_weight=0.8, kd_roughness=0.0)
    r = run(m, queries(wo, same))
    assert np.allclose(r["f_eval"], KD * 0.8 / np.pi, rtol=1e-6)
    assert np.allclose(r["pdf_eval"], same[:, 2] / np.pi, rtol=1e-6)
    r = run(m, queries(wo, opp))
    assert np.all(r["f_eval"] == 0) and np.all(r["pdf_eval"] == 0)
    m = surface(kd_weight=0.8, kd_roughness=0.0, thin_walled=1)
    r = run(m, queries(wo, opp))
    assert np.allclose(r["f_eval"], KD * 0.8 / np.pi, rtol=1e-6)
    assert np.allclose(r["pdf_eval"], -opp[:, 2] / np.pi, rtol=1e-6)
    r = run(m, queries(wo, same))
    assert np.all(r["f_eval"] == 0) and np.all(r["pdf_eval"] == 0)


@pytest.mark.parametrize("rough", [0.2, 0.5, 1.0])
def test_oren_nayar_view_along_normal(rough):
    rng = np.random.default_rng(1)
    wi = random_dirs(rng, 200)
    wi[:, 2] = np.abs(wi[:, 2])
    r = run(surface(kd_weight=0.7, kd_roughness=rough), queries(np.tile([0, 0, 1.0], (200, 1)), wi))
    s = np.radians(rough * 90.0)
    A = 1 - s * s / (2 * (s * s + 0.33))
    assert np.allclose(r["f_eval"], KD * 0.7 * A / np.pi, rtol=1e-5)


@pytest.mark.parametrize("name", ["oren", "conductor", "fblend_dielectric", "fblend_metallic", "thinfilm", "sheen",
                                  "coat", "oren+conductor+thinfilm+coat", "fblend+thinfilm+sheen+coat"])
def test_reciprocity(name):
    rng = np.random.default_rng(2)
    wo, wi, _ = _same_and_opposite(rng, 2000)
    m = material(name)
    a = run(m, queries(wo, wi))["f_eval"].astype(np.float64)
    b = run(m, queries(wi, wo))["f_eval"].astype(np.float64)
    assert np.all(a >= 0) and np.any(a > 0)
    assert np.allclose(a, b, rtol=2e-5, atol=1e-12), np.max(np.abs(a - b) / np.maximum(b, 1e-12))


def _thin_film_normal(nf, nm, d, rough=0.5):
    m = surface(ks_roughness=rough, thin_film_thickness=d, thin_film_ior=nf, ks_ior=nm)
    z = np.array([[0, 0, 1.0]])
    r = run(m, queries(z, z))
    a = rough * rough
    return r["f_eval"][0].astype(np.float64) * 4 * np.pi * a * a


def test_thin_film_index_matched_is_single_interface():
    for n in (1.3, 1.5, 2.2):
        F = _thin_film_normal(n, n, 380.0)
        assert np.allclose(F, ((n - 1) / (n + 1)) ** 2, rtol=1e-5), (n, F)


@pytest.mark.parametrize("nf,nm,d", [(1.33, 1.5, 320.0), (1.2, 1.8, 500.0), (1.45, 2.4, 150.0)])
def test_thin_film_airy_at_normal_incidence(nf, nm, d):
    F = _thin_film_normal(nf, nm, d)
    r01 = (1 - nf) / (1 + nf)
    r12 = (nf - nm) / (nf + nm)
    c = np.cos(4 * np.pi * nf * d / LAM)
    R = (r01 ** 2 + r12 ** 2 + 2 * r01 * r12 * c) / (1 + r01 ** 2 * r12 ** 2 + 2 * r01 * r12 * c)
    assert np.allclose(F, R, rtol=2e-5), (F, R)
    assert np.ptp(F) > 0.01          # the lanes see different phases


@pytest.mark.parametrize("gloss,ior,w", [(0.0, 1.5, 0.8), (0.6, 1.4, 0.5), (1.0, 2.0, 1.0)])
def test_clearcoat_along_normal(gloss, ior, w):
    m = surface(coat_weight=w, coat_gloss=gloss, coat_ior=ior)
    z = np.array([[0, 0, 1.0]])
    f = run(m, queries(z, z))["f_eval"][0].astype(np.float64)
    a = (1 - gloss) * 0.1 + gloss * 0.001
    D = (a * a - 1) / (np.pi * np.log(a * a) * a * a)
    Fr = ((ior - 1) / (ior + 1)) ** 2
    assert np.allclose(f, COAT * w * Fr * D / 16, rtol=1e-5), (f, COAT * w * Fr * D / 16)


def _pdf_mass_normal(m, nt=40000):
    th = (np.arange(nt) + 0.5) * (np.pi / 2) / nt
    wi = np.stack([np.sin(th), np.zeros(nt), np.cos(th)], -1)
    r = run(m, queries(np.tile([0, 0, 1.0], (nt, 1)), wi))
    return float(np.sum(r["pdf_eval"].astype(np.float64) * 2 * np.pi * np.sin(th)) * (np.pi / 2) / nt)


@pytest.mark.parametrize("rough", [0.25, 0.45, 0.8])
def test_sheen_pdf_mass(rough):
    mass = _pdf_mass_normal(surface(sheen_weight=1.0, sheen_roughness=rough))
    want = (1 / np.sqrt(2)) ** (1 / rough + 2)
    assert abs(mass - want) < 2e-4 * max(want, 1e-3), (mass, want)


@pytest.mark.parametrize("gloss", [0.0, 0.3])
def test_clearcoat_pdf_mass(gloss):
    a = (1 - gloss) * 0.1 + gloss * 0.001
    mass = _pdf_mass_normal(surface(coat_weight=1.0, coat_gloss=gloss))
    want = 1 - np.log((1 + a * a) / 2) / np.log(a * a)
    assert abs(mass - want) < 1e-3, (mass, want)


WOS = unit_rows(np.array([[0.0, 0.0, 1.0], [0.45, 0.2, 0.87], [0.9, -0.3, 0.3]]))


def _albedo_quadrature(m, wo, nt=900, nph=900):
    d, w = sphere_grid(nt, nph)
    r = run(m, queries(np.tile(wo, (d.shape[0], 1)), d))
    return (r["f_eval"].astype(np.float64) * (np.abs(d[:, 2]) * w)[:, None]).sum(0)


def _albedo_sampled(m, wo, n=400, density=None):
    u = strata(n)
    r = run(m, queries(np.tile(wo, (u.shape[0], 1)), u=u, vndf=True))
    ok = r["pdf_sample"] > 0
    wi = r["wi_sample"].astype(np.float64)
    pdf = r["pdf_sample"].astype(np.float64) if density is None else density(wo, wi)
    g = np.zeros((u.shape[0], 4))
    g[ok] = r["f_sample"][ok] * (np.abs(wi[ok, 2]) / pdf[ok])[:, None]
    return g.mean(0)


@pytest.mark.parametrize("name", MATCHED)
def test_sampling_matches_pdf_and_evaluate(name):
    m = material(name)
    for wo in WOS:
        A = _albedo_sampled(m, wo)
        B = _albedo_quadrature(m, wo)
        assert np.all(B > 1e-3)
        assert np.allclose(A, B, rtol=4e-3), (name, wo, A, B)


def test_sheen_samples_cosine_half_vectors():
    """The sheen sampler's true density of wi (cosine-weighted wh over the
    reflection Jacobian 1 / (4 wo.wh)), not its reported pdf, turns its
    samples into an unbiased estimate of the albedo."""
    m = material("sheen")

    def true_density(wo, wi):
        wh = unit_rows(wi + wo)
        return np.abs(wh[:, 2]) / np.pi / (4 * np.abs(wh @ wo))

    for wo in WOS:
        A = _albedo_sampled(m, wo, density=true_density)
        B = _albedo_quadrature(m, wo)
        assert np.allclose(A, B, rtol=4e-3), (wo, A, B)


@pytest.mark.parametrize("name", ["oren+conductor+thinfilm+coat", "fblend+thinfilm+sheen+coat", "fblend_metallic"])
def test_sample_equals_pdf_and_evaluate_at_sample(name):
    rng = np.random.default_rng(5)
    n = 3000
    wo = random_dirs(rng, n)
    m = material(name)
    r = run(m, queries(wo, u=rng.random((n, 2))))
    ok = r["pdf_sample"] > 0
    assert ok.mean() > 0.5
    e = run(m, queries(wo[ok], r["wi_sample"][ok].astype(np.float64)))
    # wi_sample comes back rounded to fp32: near grazing that moves pdf and f
    # by up to a few 1e-5
    assert np.allclose(r["pdf_sample"][ok], e["pdf_eval"], rtol=1e-4)
    assert np.allclose(r["f_sample"][ok], e["f_eval"], rtol=1e-4, atol=1e-12)


def test_specular_lobe_counts_in_sample_pdf():
    rng = np.random.default_rng(6)
    n = 3000
    wo = random_dirs(rng, n)
    m = material("difftrans+fspec+coat")
    r = run(m, queries(wo, u=rng.random((n, 2))))
    spec = (r["sampled_type"] & flgpu.BSDF_SPECULAR) != 0
    ok = (r["pdf_sample"] > 0) & ~spec
    assert 0.2 < spec.mean() < 0.45 and ok.mean() > 0.3
    e = run(m, queries(wo[ok], r["wi_sample"][ok].astype(np.float64)))
    # sample: (difftrans + coat) / 3 components; pdf without specular: / 2
    assert np.allclose(r["pdf_sample"][ok], e["pdf_eval"] * 2 / 3, rtol=1e-4)
    assert np.allclose(r["f_sample"][ok], e["f_eval"], rtol=1e-4, atol=1e-12)


def test_fp32_oracle_agrees_with_fp64():
    rng = np.random.default_rng(7)
    n = 4000
    wo, wi = random_dirs(rng, n), random_dirs(rng, n)
    m = material("fblend+thinfilm+sheen+coat")
    q = queries(wo, wi, u=rng.random((n, 2)))
    a = oracle.bsdf_standard(m, q, 64)
    b = oracle.bsdf_standard(m, q, 32)
    assert np.allclose(a["f_eval"], b["f_eval"], rtol=1e-3, atol=1e-7)
    assert np.allclose(a["pdf_eval"], b["pdf_eval"], rtol=1e-3, atol=1e-7)
