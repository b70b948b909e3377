// fl_shade.h -- hero-wavelength spectra and the BSDFs of the render path.
//
// Each function restates the reference routine named beside it, in fp32 and in
// the same operation order, but over 4 hero-wavelength lanes instead of the
// reference's 60 fixed bins (color/principalspectrum.h:7-9):
//   BSDF frame / evaluate / sample / pdf   shading/bsdf.cpp:6-151
//   LambertBRDF                            shading/bxdfs/lambert.cpp:6-25
//   cosine_hemisphere_sample               core/shadingcore.h:65-93
//   MicrofacetReflectionBRDF               shading/bxdfs/microfacetreflection.cpp:13-61
//   MicrofacetDistribution G / pdf         shading/microfacets/microfacet.cpp:6-16
//   GGXDistribution D / lambda / sample_wh shading/microfacets/ggx.cpp:7-57
//   FresnelConductor                       shading/fresnel/fresnelconductor.cpp:9-33
//   FresnelDielectric (pure function)      shading/fresnel/fresneldielectric.cpp:8-38
//   FresnelSpecularBXDF                    shading/closures/fresnelspecular.cpp:8-53
#pragma once
#include "fl_math.h"

namespace fl {

// BxDFType bits, shading/bxdfs/bxdf.h:8-15
enum : uint32_t {
    BSDF_REFLECTION = 1u << 0,
    BSDF_TRANSMISSION = 1u << 1,
    BSDF_DIFFUSE = 1u << 2,
    BSDF_GLOSSY = 1u << 3,
    BSDF_SPECULAR = 1u << 4,
};

enum : int { MAT_NONE = 0, MAT_LAMBERT = 1, MAT_CONDUCTOR = 2, MAT_DIELECTRIC = 3 };

// ---------------------------------------------------------------- spectra --
// A wavelength set: 4 lanes, lane k = lo + (l0 - lo + k*(hi-lo)/4) mod (hi-lo).
// Parity grid: the reference's 60 bins of 5 nm over 400..700, piecewise
// constant.  Hero grid: 471 samples at 1 nm over 360..830, linear.
struct SpecGrid {
    float lo, hi;        // sampling range
    float g0, ginv;      // grid origin and inverse spacing
    int last;            // last valid interpolation base index
    int linear;          // 1: interpolate, 0: piecewise constant
    int stride;          // floats per table
};

struct Lanes {
    float lam[4];
    int idx[4];
    float frac[4];
};

FL_HD Lanes make_lanes(const SpecGrid &g, float lam0) {
    Lanes L;
    float span = g.hi - g.lo;
    for (int k = 0; k < 4; ++k) {
        float l = lam0 + (float)k * (span * 0.25f);
        if (l >= g.hi) l -= span;
        L.lam[k] = l;
        float x = (l - g.g0) * g.ginv;
        int i = (int)x;
        i = i < 0 ? 0 : (i > g.last ? g.last : i);
        L.idx[k] = i;
        L.frac[k] = g.linear ? (x - (float)i) : 0.f;
    }
    return L;
}

FL_HD float spec_at(const float *tab, const Lanes &L, int k) {
    float a = tab[L.idx[k]];
    float b = tab[L.idx[k] + 1];
    return a + L.frac[k] * (b - a);
}

// ------------------------------------------------------------ local trig --
// shadingcore.h:10-55
FL_HD float cos2t(v3 w) { return w.z * w.z; }
FL_HD float sin2t(v3 w) { return fmaxf(0.f, 1.f - cos2t(w)); }
FL_HD float sint(v3 w) { return sqrtf(sin2t(w)); }
FL_HD float tant(v3 w) { return sint(w) / w.z; }
FL_HD float tan2t(v3 w) { return sin2t(w) / cos2t(w); }
FL_HD float cosphi(v3 w) { float s = sint(w); return (s == 0.f) ? 1.f : clampf(w.x / s, -1.f, 1.f); }
FL_HD float sinphi(v3 w) { float s = sint(w); return (s == 0.f) ? 0.f : clampf(w.y / s, -1.f, 1.f); }
FL_HD float cos2phi(v3 w) { float c = cosphi(w); return c * c; }
FL_HD float sin2phi(v3 w) { float s = sinphi(w); return s * s; }
FL_HD bool same_hemi(v3 a, v3 b) { return (a.z * b.z) > 0.f; }

// concentric_sample_disk + cosine_hemisphere_sample, shadingcore.h:65-89.
// z = sqrt(1 - dx^2 - dy^2) with dx^2 + dy^2 = r^2 and r = 2u - 1 (u the
// coordinate that set r) equals sqrt(4u(1-u)): that form keeps full relative
// precision for grazing samples, where 1 - dx^2 - dy^2 cancels.
FL_HD v3 cosine_hemisphere(float u0, float u1) {
    float ox = u0 * 2.f - 1.f, oy = u1 * 2.f - 1.f;
    float dx = 0.f, dy = 0.f, ur = 0.5f;
    if (!(ox == 0.f && oy == 0.f)) {
        float theta, r;
        if (fabsf(ox) > fabsf(oy)) {
            r = ox;
            ur = u0;
            theta = kPiOver4 * (oy / ox);
        } else {
            r = oy;
            ur = u1;
            theta = kPiOver2 - kPiOver4 * (ox / oy);
        }
        dx = cosf(theta) * r;
        dy = sinf(theta) * r;
    }
    float z = 2.f * sqrtf(fmaxf(0.f, ur * (1.f - ur)));
    return v3(dx, dy, z);
}

// --------------------------------------------------------------- GGX -----
struct GGX {
    float ax, ay;
};

// GGXDistribution::evaluate_distribution, ggx.cpp:13-21.  For a unit vector
// (cos^2 phi/ax^2 + sin^2 phi/ay^2) tan^2 theta = (x^2/ax^2 + y^2/ay^2)/z^2;
// that form avoids 1 - cos^2 theta, which cancels near the pole where small
// alphas amplify the error by 1/alpha^2.  z = 0 (tan = inf) gives 0 as in the
// reference.
FL_HD float ggx_D(const GGX &g, v3 wh) {
    float z2 = wh.z * wh.z;
    if (z2 == 0.f) return 0.f;
    float e = (wh.x * wh.x / (g.ax * g.ax) + wh.y * wh.y / (g.ay * g.ay)) / z2;
    return 1.f / (kPi * g.ax * g.ay * (z2 * z2) * (1.f + e) * (1.f + e));
}

// GGXDistribution::lambda, ggx.cpp:23-30, with alpha^2 tan^2 theta =
// (x^2 ax^2 + y^2 ay^2)/z^2 and (-1 + sqrt(1+a))/2 = a / (2 (1 + sqrt(1+a))).
FL_HD float ggx_lambda(const GGX &g, v3 w) {
    float z2 = w.z * w.z;
    if (z2 == 0.f) return 0.f;
    float a2t2 = (w.x * w.x * g.ax * g.ax + w.y * w.y * g.ay * g.ay) / z2;
    return a2t2 / (2.f * (1.f + sqrtf(1.f + a2t2)));
}

FL_HD float ggx_G(const GGX &g, v3 wo, v3 wi) { return 1.f / (1.f + ggx_lambda(g, wo) + ggx_lambda(g, wi)); }
FL_HD float ggx_G1(const GGX &g, v3 w) { return 1.f / (1.f + ggx_lambda(g, w)); }

// MicrofacetDistribution::pdf = D(wh)*|cos(wh)|, microfacet.cpp:14-16.
// With vndf=true: the density sample_wh actually draws from (Heitz 2018).
FL_HD float ggx_pdf_wh(const GGX &g, v3 wo, v3 wh, bool vndf) {
    if (!vndf) return ggx_D(g, wh) * fabsf(wh.z);
    return ggx_D(g, wh) * ggx_G1(g, wo) * fabsf(dot(wo, wh)) / fabsf(wo.z);
}

// GGXDistribution::sample_wh, ggx.cpp:32-57 (visible-normal sampling),
// evaluated in fp64: 1 - t1^2 - t2^2 cancels for samples near the rim of the
// projected disk, and the small-alpha stretch amplifies that error, so fp32
// cannot stay within 1e-5 of the exact direction there.
FL_HD v3 ggx_sample_wh(const GGX &g, v3 wo, float u0, float u1) {
    const double ax = g.ax, ay = g.ay;
    double vx = wo.x * ax, vy = wo.y * ay, vz = wo.z;
    double vl = sqrt(vx * vx + vy * vy + vz * vz);
    vx /= vl; vy /= vl; vz /= vl;
    if (vz < 0.0) { vx = -vx; vy = -vy; vz = -vz; }
    double lensq = vx * vx + vy * vy;
    double T1x = 1.0, T1y = 0.0;
    if (lensq > 0.0) { double il = sqrt(lensq); T1x = -vy / il; T1y = vx / il; }
    // T2 = cross(vh, T1), T1.z = 0
    double T2x = vy * 0.0 - vz * T1y, T2y = vz * T1x - vx * 0.0, T2z = vx * T1y - vy * T1x;
    double r = sqrt((double)u0);
    double phi = 2.0 * 3.14159265358979323846 * (double)u1;
    double t1 = r * cos(phi);
    double t2 = r * sin(phi);
    double s = (1.0 + vz) * 0.5;
    t2 = (1.0 - s) * sqrt(1.0 - t1 * t1) + s * t2;
    double nz = sqrt(fmax(0.0, 1.0 - t1 * t1 - t2 * t2));
    double Nx = T1x * t1 + T2x * t2 + vx * nz;
    double Ny = T1y * t1 + T2y * t2 + vy * nz;
    double Nz = 0.0 * t1 + T2z * t2 + vz * nz;
    double hx = ax * Nx, hy = ay * Ny, hz = fmax((double)1e-6f, Nz);
    double hl = sqrt(hx * hx + hy * hy + hz * hz);
    return v3((float)(hx / hl), (float)(hy / hl), (float)(hz / hl));
}

// FresnelConductor::evaluate with etaI = 1, fresnelconductor.cpp:9-33
FL_HD float fresnel_conductor(float cosI, float eta, float k) {
    cosI = fabsf(cosI);
    cosI = clampf(cosI, -1.f, 1.f);
    float c2 = cosI * cosI;
    float s2 = 1.f - c2;
    float eta2 = eta * eta, k2 = k * k;
    float t0 = eta2 - k2 - s2;
    float a2b2 = sqrtf(t0 * t0 + 4.f * eta2 * k2);
    float t1 = a2b2 + c2;
    float a = sqrtf(0.5f * (a2b2 + t0));
    float t2 = 2.f * cosI * a;
    float Rs = (t1 - t2) / (t1 + t2);
    float t3 = c2 * a2b2 + s2 * s2;
    float t4 = t2 * s2;
    float Rp = Rs * (t3 - t4) / (t3 + t4);
    return 0.5f * (Rp + Rs);
}

// FresnelDielectric::evaluate (swap kept local), fresneldielectric.cpp:8-38
FL_HD float fresnel_dielectric(float cosI, float etaI, float etaT) {
    cosI = clampf(cosI, -1.f, 1.f);
    if (!(cosI > 0.f)) {
        float t = etaI; etaI = etaT; etaT = t;
        cosI = fabsf(cosI);
    }
    float sinI = sqrtf(fmaxf(0.f, 1.f - cosI * cosI));
    float sinT = etaI / etaT * sinI;
    if (sinT >= 1.f) return 1.f;
    float cosT = sqrtf(fmaxf(0.f, 1.f - sinT * sinT));
    float Rparl = (etaT * cosI - etaI * cosT) / (etaT * cosI + etaI * cosT);
    float Rperp = (etaI * cosI - etaT * cosT) / (etaI * cosI + etaT * cosT);
    return (Rparl * Rparl + Rperp * Rperp) * 0.5f;
}

// fresnel_dielectric in fp64 (same steps)
FL_HD double fresnel_dielectric_d(double cosI, double etaI, double etaT) {
    cosI = cosI < -1.0 ? -1.0 : (cosI > 1.0 ? 1.0 : cosI);
    if (!(cosI > 0.0)) {
        double t = etaI; etaI = etaT; etaT = t;
        cosI = fabs(cosI);
    }
    double sinI = sqrt(fmax(0.0, 1.0 - cosI * cosI));
    double sinT = etaI / etaT * sinI;
    if (sinT >= 1.0) return 1.0;
    double cosT = sqrt(fmax(0.0, 1.0 - sinT * sinT));
    double Rparl = (etaT * cosI - etaI * cosT) / (etaT * cosI + etaI * cosT);
    double Rperp = (etaI * cosI - etaT * cosT) / (etaI * cosI + etaT * cosT);
    return (Rparl * Rparl + Rperp * Rperp) * 0.5;
}

// ----------------------------------------------------------------- BSDF ----
// One lobe per material in every supported material (bsdf.h:41 allows 8).
struct Bsdf {
    int kind;
    GGX ggx;
    float weight;        // MicrofacetReflection weight (ksweight)
    float s1[4], s2[4];  // per lane: Lambert (R, -), conductor (eta, k), dielectric (R, T)
    float ior;           // dielectric: eta at the hero wavelength (lane 0)
    bool dispersive;     // dielectric: eta depends on wavelength (secondary lanes end)
    bool vndf_pdf;
};

FL_HD uint32_t lobe_type(int kind) {
    switch (kind) {
    case MAT_LAMBERT: return BSDF_REFLECTION | BSDF_DIFFUSE;
    case MAT_CONDUCTOR: return BSDF_REFLECTION | BSDF_GLOSSY;
    case MAT_DIELECTRIC: return BSDF_REFLECTION | BSDF_TRANSMISSION | BSDF_SPECULAR;
    }
    return 0;
}

FL_HD bool is_specular_kind(int kind) { return kind == MAT_DIELECTRIC; }

// MicrofacetReflectionBRDF::evaluate, microfacetreflection.cpp:13-32 (R = 1).
// `whp` (optional) is the half vector when it is already known exactly (a
// sampled microfacet normal): normalize(wo + wi) cancels for grazing pairs.
FL_HD void microfacet_f(const Bsdf &b, v3 wo, v3 wi, float f[4], const v3 *whp = nullptr) {
    float cO = fabsf(wo.z), cI = fabsf(wi.z);
    v3 wh = wi + wo;
    for (int k = 0; k < 4; ++k) f[k] = 0.f;
    if (cI == 0.f || cO == 0.f) return;
    if (wh.x == 0.f && wh.y == 0.f && wh.z == 0.f) return;
    wh = whp ? *whp : normalize(wh);
    float cosF = dot(wi, face_forward(wh, v3(0.f, 0.f, 1.f)));
    float D = ggx_D(b.ggx, wh);
    float G = ggx_G(b.ggx, wo, wi);
    for (int k = 0; k < 4; ++k) {
        float F = fresnel_conductor(cosF, b.s1[k], b.s2[k]);
        f[k] = 1.f * b.weight * D * G * F / (4.f * cO * cI);
    }
}

// BxDF::pdf for the two non-specular lobes
FL_HD float lobe_pdf(const Bsdf &b, v3 wo, v3 wi) {
    if (b.kind == MAT_LAMBERT) return same_hemi(wo, wi) ? fabsf(wi.z) * kInvPi : 0.f;
    if (b.kind == MAT_CONDUCTOR) {
        if (!same_hemi(wo, wi)) return 0.f;
        v3 wh = normalize(wo + wi);
        return ggx_pdf_wh(b.ggx, wo, wh, b.vndf_pdf) / (4.f * dot(wo, wh));
    }
    return 0.f;
}

// BSDF::evaluate(wo, wi, BSDF_ALL & ~BSDF_SPECULAR) in the local frame;
// `reflect` is dot(wiW,ng)*dot(woW,ng) > 0 computed by the caller in world space.
FL_HD void bsdf_eval_nonspec(const Bsdf &b, v3 wo, v3 wi, bool reflect, float f[4]) {
    for (int k = 0; k < 4; ++k) f[k] = 0.f;
    if (wo.z == 0.f) return;
    if (!reflect) return;  // both non-specular lobes are reflection-only
    if (b.kind == MAT_LAMBERT) {
        for (int k = 0; k < 4; ++k) f[k] = b.s1[k] * kInvPi * 1.f;
    } else if (b.kind == MAT_CONDUCTOR) {
        microfacet_f(b, wo, wi, f);
    }
}

// BSDF::pdf(wo, wi, BSDF_ALL & ~BSDF_SPECULAR)
FL_HD float bsdf_pdf_nonspec(const Bsdf &b, v3 wo, v3 wi) {
    if (b.kind != MAT_LAMBERT && b.kind != MAT_CONDUCTOR) return 0.f;
    if (wo.z == 0.f) return 0.f;
    return lobe_pdf(b, wo, wi);
}

// BSDF::sample(wo, u, BSDF_ALL) in the local frame.  Returns the sampled
// type (0 on failure).  *reflect_of(wi) must be supplied by the caller via
// the world-space test after the lobe sample; for the local-frame query path
// the caller passes ng = (0,0,1) semantics, see bsdf_sample_local.
struct BsdfSample {
    v3 wi;
    v3 wh;               // sampled microfacet normal (conductor)
    float f[4];
    float pdf;
    uint32_t type;
    bool secondary_terminated;
};

// The lobe-level sample (BxDF::sample).  Returns f before BSDF::sample's
// non-specular re-evaluation.
FL_HD void lobe_sample(const Bsdf &b, v3 wo, float u0, float u1, BsdfSample &s) {
    s.pdf = 0.f;
    s.type = lobe_type(b.kind);
    s.secondary_terminated = false;
    for (int k = 0; k < 4; ++k) s.f[k] = 0.f;
    if (b.kind == MAT_LAMBERT) {
        // LambertBRDF::sample, lambert.cpp:10-20
        v3 wi = cosine_hemisphere(u0, u1);
        if (wo.z < 0.f) wi.z *= -1.f;
        s.wi = wi;
        s.pdf = same_hemi(wo, wi) ? fabsf(wi.z) * kInvPi : 0.f;
        for (int k = 0; k < 4; ++k) s.f[k] = b.s1[k] * kInvPi * 1.f;
    } else if (b.kind == MAT_CONDUCTOR) {
        // MicrofacetReflectionBRDF::sample, microfacetreflection.cpp:34-53
        if (wo.z == 0.f) return;
        v3 wh = ggx_sample_wh(b.ggx, wo, u0, u1);
        if (dot(wo, wh) < 0.f) return;
        v3 wi = reflect(wo, wh);
        if (!same_hemi(wo, wi)) return;
        s.wi = wi;
        s.wh = wh;
        s.pdf = ggx_pdf_wh(b.ggx, wo, wh, b.vndf_pdf) / (4.f * dot(wo, wh));
        microfacet_f(b, wo, wi, s.f, &wh);
    } else if (b.kind == MAT_DIELECTRIC) {
        // FresnelSpecularBXDF::sample, fresnelspecular.cpp:21-48; the hero lane
        // (lane 0) decides; a wavelength-dependent eta terminates lanes 1..3.
        // F, 1-F and the refracted direction are formed in fp64: 1-F and cos(theta_t)
        // cancel near grazing incidence / the critical angle.
        float ior = b.ior;
        bool dispersive = b.dispersive;
        double Fd = fresnel_dielectric_d(wo.z, 1.0, ior);
        float F = (float)Fd;
        if (u0 < F) {
            v3 wi = v3(-wo.x, -wo.y, wo.z);
            s.type = BSDF_SPECULAR | BSDF_REFLECTION;
            s.wi = wi;
            s.pdf = F;
            for (int k = 0; k < 4; ++k) s.f[k] = b.s1[k] * F / fabsf(wi.z);
        } else {
            bool entering = wo.z > 0.f;
            double etaI = entering ? 1.0 : (double)ior;
            double etaT = entering ? (double)ior : 1.0;
            double eta = etaI / etaT;
            double nz = entering ? 1.0 : -1.0;   // face_forward((0,0,1), wo)
            double cosI = (double)wo.z * nz;
            double sinT2 = eta * eta * fmax(0.0, 1.0 - cosI * cosI);
            if (sinT2 >= 1.0) return;
            double cosT = sqrt(1.0 - sinT2);
            double k2 = eta * cosI - cosT;
            v3 wi((float)(-eta * wo.x), (float)(-eta * wo.y), (float)(-eta * (double)wo.z + nz * k2));
            double T = 1.0 - Fd;
            double sc = (etaI * etaI) / (etaT * etaT) / fabs(-eta * (double)wo.z + nz * k2);
            s.type = BSDF_SPECULAR | BSDF_TRANSMISSION;
            s.wi = wi;
            s.pdf = (float)T;
            for (int k = 0; k < 4; ++k) s.f[k] = (float)((double)b.s2[k] * T * sc);
        }
        s.secondary_terminated = dispersive;
    }
}

// BSDF::sample's u remap for one matching component, bsdf.cpp:63-75:
// comp = min(floor(u.x*1), 0) = 0, u.x' = min(u.x*1 - 0, 1 - SHADOW_EPSILON).
FL_HD float remap_u0(float u0) { return fminf(u0 * 1.f - 0.f, 1.f - kShadowEps); }

} // namespace fl
