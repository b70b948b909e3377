// fl_spectra.cpp -- per-scene spectral tables for the two wavelength grids.
//
// Parity grid: the reference's 60 bins over 400..700 nm, computed exactly as
// the reference does (float arithmetic in the same order):
//   average_spectrum_samples      color/spectrumoperations.cpp:28-56
//   bin edges (lerp)              color/spectruminits.cpp:8-10, from_spd :88-103
//   from_reflectance_rgb          color/spectrumoperations.cpp:113-147
//   metal eta/k bin averages      color/metallurgy.cpp:7-40
// Hero grid: 471 samples at 1 nm over 360..830 nm: the same data evaluated
// pointwise (linear interpolation, flat extrapolation), so that a hero
// wavelength sees the continuous spectrum.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "../data/fl_tables.inc"
#include "fl_internal.h"

namespace {

inline float lerpf(float t, float v1, float v2) { return (1.f - t) * v1 + t * v2; }

float average_spectrum_samples(const float *lam, const float *vals, int n, float l0, float l1) {
    if (l1 <= lam[0]) return vals[0];
    if (l0 >= lam[n - 1]) return vals[n - 1];
    if (n == 1) return vals[0];
    float sum = 0;
    if (l0 < lam[0]) sum += vals[0] * (lam[0] - l0);
    if (l1 > lam[n - 1]) sum += vals[n - 1] * (l1 - lam[n - 1]);
    int i = 0;
    while (l0 > lam[i + 1]) ++i;
    auto interp = [&](float w, int j) { return lerpf((w - lam[j]) / (lam[j + 1] - lam[j]), vals[j], vals[j + 1]); };
    for (; i + 1 < n && l1 >= lam[i]; ++i) {
        float s0 = std::max(l0, lam[i]);
        float s1 = std::min(l1, lam[i + 1]);
        sum += 0.5 * (interp(s0, i) + interp(s1, i)) * (s1 - s0);
    }
    return sum / (l1 - l0);
}

// pointwise value with flat extrapolation
float sample_at(const float *lam, const float *vals, int n, float l) {
    if (l <= lam[0]) return vals[0];
    if (l >= lam[n - 1]) return vals[n - 1];
    int i = 0;
    while (l > lam[i + 1]) ++i;
    return lerpf((l - lam[i]) / (lam[i + 1] - lam[i]), vals[i], vals[i + 1]);
}

constexpr int kBins = 60;
constexpr int kHeroN = 471;

void bin_edges(int i, float *wl0, float *wl1) {
    *wl0 = lerpf(float(i) / float(kBins), 400.f, 700.f);
    *wl1 = lerpf(float(i + 1) / float(kBins), 400.f, 700.f);
}

// A sampled function on the grid of `mode`.
void project(int mode, const float *lam, const float *vals, int n, float *out) {
    if (mode == FL_SPECTRAL_PARITY) {
        for (int i = 0; i < kBins; ++i) {
            float a, b;
            bin_edges(i, &a, &b);
            out[i] = average_spectrum_samples(lam, vals, n, a, b);
        }
        out[kBins] = out[kBins - 1];
    } else {
        for (int i = 0; i < kHeroN; ++i) out[i] = sample_at(lam, vals, n, 360.f + (float)i);
        out[kHeroN] = out[kHeroN - 1];
    }
}

struct Basis {
    std::vector<float> white, cyan, magenta, yellow, red, green, blue;
};

Basis refl_basis(int mode, int N) {
    Basis b;
    auto mk = [&](const float *v, std::vector<float> &o) {
        o.assign(N + 1, 0.f);
        project(mode, FL_RGB2SPEC_LAMBDA, v, FL_RGB2SPEC_N, o.data());
    };
    mk(FL_RGBREFL_WHITE, b.white);
    mk(FL_RGBREFL_CYAN, b.cyan);
    mk(FL_RGBREFL_MAGENTA, b.magenta);
    mk(FL_RGBREFL_YELLOW, b.yellow);
    mk(FL_RGBREFL_RED, b.red);
    mk(FL_RGBREFL_GREEN, b.green);
    mk(FL_RGBREFL_BLUE, b.blue);
    return b;
}

// from_reflectance_rgb on the grid: r accumulates `coef * basis` per entry
// (BaseSpectrum operator+= / operator*(float)), then *= .94 and clamp >= 0.
void from_reflectance_rgb(const Basis &B, const float rgb[3], int N, float *out) {
    std::vector<float> r(N, 0.f);
    auto add = [&](float c, const std::vector<float> &s) {
        for (int i = 0; i < N; ++i) r[i] += s[i] * c;
    };
    if (rgb[0] <= rgb[1] && rgb[0] <= rgb[2]) {
        add(rgb[0], B.white);
        if (rgb[1] <= rgb[2]) { add(rgb[1] - rgb[0], B.cyan); add(rgb[2] - rgb[1], B.blue); }
        else { add(rgb[2] - rgb[0], B.cyan); add(rgb[1] - rgb[2], B.green); }
    } else if (rgb[1] <= rgb[0] && rgb[1] <= rgb[2]) {
        add(rgb[1], B.white);
        if (rgb[0] <= rgb[2]) { add(rgb[0] - rgb[1], B.magenta); add(rgb[2] - rgb[0], B.blue); }
        else { add(rgb[2] - rgb[1], B.magenta); add(rgb[0] - rgb[2], B.red); }
    } else {
        add(rgb[2], B.white);
        if (rgb[0] <= rgb[1]) { add(rgb[0] - rgb[2], B.yellow); add(rgb[1] - rgb[0], B.green); }
        else { add(rgb[1] - rgb[2], B.yellow); add(rgb[0] - rgb[1], B.red); }
    }
    for (int i = 0; i < N; ++i) {
        float v = r[i] * (float).94;
        out[i] = v < 0.f ? 0.f : v;   // spectrum_clamp
    }
    out[N] = out[N - 1];
}

} // namespace

namespace flx {

void build_spectral_tables(fl_scene_s *s, const fl_scene_desc *desc, int mode, std::vector<DevMaterial> &mats) {
    SpecTables &T = s->spec[mode];
    const int N = mode == FL_SPECTRAL_PARITY ? kBins : kHeroN;
    const int stride = N + 1;
    fl::SpecGrid &g = T.grid;
    if (mode == FL_SPECTRAL_PARITY) {
        g.lo = 400.f; g.hi = 700.f; g.g0 = 400.f; g.ginv = 0.2f; g.last = kBins - 1; g.linear = 0;
    } else {
        g.lo = 360.f; g.hi = 830.f; g.g0 = 360.f; g.ginv = 1.f; g.last = kHeroN - 2; g.linear = 1;
    }
    g.stride = stride;
    const uint32_t nm = desc->n_materials, ne = desc->n_emission, nstd = desc->n_standard_surfaces;
    // + the RGB->spectrum basis (normal AOV) + 7 spectra per standard surface
    const size_t slots = 3 + 2 * (size_t)nm + ne + 7 + (size_t)fl::SS_N * nstd;
    T.host.assign(slots * stride, 0.f);
    float *tab = T.host.data();
    // colour matching functions
    if (mode == FL_SPECTRAL_PARITY) {
        float cie_lam[FL_CIE_N];
        for (int i = 0; i < FL_CIE_N; ++i) cie_lam[i] = 360.f + (float)i;
        project(mode, cie_lam, FL_CIE_X, FL_CIE_N, tab + 0 * stride);
        project(mode, cie_lam, FL_CIE_Y, FL_CIE_N, tab + 1 * stride);
        project(mode, cie_lam, FL_CIE_Z, FL_CIE_N, tab + 2 * stride);
    } else {
        for (int i = 0; i < kHeroN; ++i) {
            tab[0 * stride + i] = FL_CIE_X[i];
            tab[1 * stride + i] = FL_CIE_Y[i];
            tab[2 * stride + i] = FL_CIE_Z[i];
        }
        for (int c = 0; c < 3; ++c) tab[c * stride + kHeroN] = tab[c * stride + kHeroN - 1];
    }
    Basis B = refl_basis(mode, N);
    float spd_lam[FL_SPD_N];
    for (int i = 0; i < FL_SPD_N; ++i) spd_lam[i] = 360.f + (float)i;
    mats.resize(nm);
    for (uint32_t m = 0; m < nm; ++m) {
        const fl_material &fm = desc->materials[m];
        DevMaterial &d = mats[m];
        d.kind = fm.kind;
        d.ax = d.ay = 1.f;
        d.weight = 1.f;
        d.eta_bsdf = 1.f;
        d.spec_a = (uint32_t)(3 + 2 * m);
        d.spec_b = (uint32_t)(4 + 2 * m);
        d.cauchy_b = 0.f;
        d.tex = 0;
        float *ta = tab + (size_t)d.spec_a * stride, *tb = tab + (size_t)d.spec_b * stride;
        // a reflectance given as a spectrum: a row of the SPD table, clamped
        auto spd_refl = [&](uint32_t row1, float *out) {
            project(mode, spd_lam, desc->emission_spd + (size_t)(row1 - 1) * FL_SPD_N, FL_SPD_N, out);
            for (int i = 0; i < stride; ++i) out[i] = std::max(0.f, out[i]);
        };
        if (fm.kind == FL_MAT_LAMBERT) {
            if (fm.spd_r) spd_refl(fm.spd_r, ta);
            else from_reflectance_rgb(B, fm.rgb, N, ta);
            d.tex = fm.texture;   // LambertMaterial's kd texture (lambert.cpp:12-13)
        } else if (fm.kind == FL_MAT_CONDUCTOR) {
            // standardsurface.cpp:104-106, GGXDistribution ctor ggx.cpp:7-11
            float aspect = std::sqrt(1.f - fm.anisotropic * 0.9f);
            float r2 = fm.roughness * fm.roughness;
            d.ax = std::max(0.001f, std::max(.001f, r2 / aspect));
            d.ay = std::max(0.001f, std::max(.001f, r2 * aspect));
            d.weight = fm.weight;
            int mt = (fm.metal >= 0 && fm.metal < FL_N_METALS) ? fm.metal : 0;
            project(mode, FL_METAL_LAMBDA[mt], FL_METAL_ETA[mt], FL_METAL_COUNT[mt], ta);
            project(mode, FL_METAL_LAMBDA[mt], FL_METAL_K[mt], FL_METAL_COUNT[mt], tb);
        } else if (fm.kind == FL_MAT_DIELECTRIC) {
            if (fm.spd_r) spd_refl(fm.spd_r, ta);
            else from_reflectance_rgb(B, fm.rgb, N, ta);
            if (fm.spd_t) spd_refl(fm.spd_t, tb);
            else from_reflectance_rgb(B, fm.rgb_t, N, tb);
            d.eta_bsdf = fm.ior;
            d.cauchy_b = fm.cauchy_b;
        } else if (fm.kind == FL_MAT_STANDARD) {
            d.spec_a = fm.standard;                 // index of its lobe stack
            d.eta_bsdf = s->h_std[fm.standard].eta_bsdf;
        }
    }
    for (uint32_t e = 0; e < ne; ++e)
        project(mode, spd_lam, desc->emission_spd + (size_t)e * FL_SPD_N, FL_SPD_N, tab + (3 + 2 * (size_t)nm + e) * stride);
    // the reflectance basis, white cyan magenta yellow red green blue, for
    // from_reflectance_rgb on the device (NormalIntegrator)
    s->rgb_basis = (uint32_t)(3 + 2 * (size_t)nm + ne);
    const std::vector<float> *basis[7] = {&B.white, &B.cyan, &B.magenta, &B.yellow, &B.red, &B.green, &B.blue};
    for (int j = 0; j < 7; ++j)
        std::copy(basis[j]->begin(), basis[j]->begin() + stride, tab + (s->rgb_basis + (size_t)j) * stride);
    // standard surfaces: kd, ks, kt, sheen, coat (from_reflectance_rgb, then
    // spectrum_clamp as compute_scattering_functions applies it), forge eta, k
    const uint32_t std0 = s->rgb_basis + 7;
    for (uint32_t j = 0; j < nstd; ++j) {
        const fl_standard_surface &ss = desc->standard_surfaces[j];
        fl::StdParams &P = s->h_std[j];
        for (int w = 0; w < fl::SS_N; ++w) P.slot[w] = std0 + fl::SS_N * j + (uint32_t)w;
        const float *rgbs[5] = {ss.kd, ss.ks, ss.kt, ss.sheen_color, ss.coat};
        for (int w = 0; w < 5; ++w) from_reflectance_rgb(B, rgbs[w], N, tab + (size_t)P.slot[w] * stride);
        int mt = (ss.metal >= 0 && ss.metal < FL_N_METALS) ? ss.metal : 0;
        project(mode, FL_METAL_LAMBDA[mt], FL_METAL_ETA[mt], FL_METAL_COUNT[mt], tab + (size_t)P.slot[fl::SS_ETA] * stride);
        project(mode, FL_METAL_LAMBDA[mt], FL_METAL_K[mt], FL_METAL_COUNT[mt], tab + (size_t)P.slot[fl::SS_K] * stride);
    }
}

// StandardSurfaceMaterial::compute_scattering_functions
// (shading/materials/standardsurface.cpp:72-159) with constant textures: the
// lobes it adds, in order, and the constants their constructors derive.
fl::StdParams standard_params(const fl_standard_surface &ss) {
    fl::StdParams P;
    std::memset(&P, 0, sizeof(P));
    const bool metal_wl = ss.fresnel_type == FL_FRESNEL_WAVELENGTH_METALLIC;
    P.eta_bsdf = (ss.ks_weight > 0.f && ss.kt_weight > 0.f) ? ss.ks_ior : 1.f;
    P.kd_weight = ss.kd_weight;
    P.tex[fl::SS_KD] = ss.kd_texture;
    P.tex[fl::SS_KS] = ss.ks_texture;
    P.tex[fl::SS_KT] = ss.kt_texture;
    P.tex[fl::SS_SHEEN] = ss.sheen_texture;
    P.tex[fl::SS_COAT] = ss.coat_texture;
    if ((ss.kd_weight > 0.f && ss.ks_weight <= 0.f) || (ss.kd_weight > 0.f && ss.ks_weight > 0.f && metal_wl)) {
        P.lobe[P.n++] = ss.thin_walled ? fl::SL_DIFFTRANS : fl::SL_OREN;
        fl::oren_nayar_AB(ss.kd_roughness * 90.f, &P.on_A, &P.on_B);   // remap [0-1] roughness to [0-90] sigma
    }
    const float aspect = std::sqrt(1.f - ss.ks_anisotropic * 0.9f);
    const float r2 = ss.ks_roughness * ss.ks_roughness;
    P.ax = std::max(.001f, r2 / aspect);
    P.ay = std::max(.001f, r2 * aspect);
    P.ks_ior = ss.ks_ior;
    P.ro = ((1.f - ss.ks_ior) / (1.f + ss.ks_ior)) * ((1.f - ss.ks_ior) / (1.f + ss.ks_ior));
    P.ks_weight = ss.ks_weight;
    P.fresnel_type = ss.fresnel_type;
    if (ss.ks_weight > 0.f) {
        if (ss.kt_weight > 0.f) {
            P.lobe[P.n++] = fl::SL_FSPEC;
        } else if (metal_wl) {
            P.lobe[P.n++] = fl::SL_CONDUCTOR;
        } else if (ss.fresnel_type == FL_FRESNEL_DIELECTRIC || ss.fresnel_type == FL_FRESNEL_METALLIC) {
            P.lobe[P.n++] = fl::SL_FBLEND;
            fl::oren_nayar_AB(ss.kd_roughness, &P.fb_A, &P.fb_B);       // the blend's OrenNayar takes kdroughness as is
        }
    }
    if (ss.thin_film_thickness > 0.f) {
        P.lobe[P.n++] = fl::SL_THINFILM;
        P.tf_ior = ss.thin_film_ior;
        P.tf_thickness = ss.thin_film_thickness;
    }
    if (ss.sheen_weight > 0.f) {
        P.lobe[P.n++] = fl::SL_SHEEN;
        P.sheen_weight = ss.sheen_weight;
        P.sheen_rough = std::max(.001f, ss.sheen_roughness);
    }
    if (ss.coat_weight > 0.f) {
        P.lobe[P.n++] = fl::SL_COAT;
        P.coat_weight = ss.coat_weight;
        P.coat_ior = ss.coat_ior;
        P.coat_gloss = (1.f - ss.coat_gloss) * .1f + ss.coat_gloss * .001f;   // lerp(gloss, .1, .001)
    }
    return P;
}

} // namespace flx
