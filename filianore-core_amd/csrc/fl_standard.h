// fl_standard.h -- StandardSurfaceMaterial's lobe stack on the device.
//
// compute_scattering_functions (shading/materials/standardsurface.cpp:72-159)
// adds up to five lobes to the BSDF; the host resolves which (StdParams), the
// device evaluates them over the 4 wavelength lanes in fp32, in the reference's
// operation order:
//   OrenNayarBRDF                    shading/bxdfs/orennayar.cpp:6-60
//   DiffuseTransmission              shading/bxdfs/diffusetransmission.cpp:6-32
//   FresnelSpecularBXDF              (fl_shade.h, MAT_DIELECTRIC)
//   MicrofacetReflection + conductor (fl_shade.h, MAT_CONDUCTOR)
//   FresnelBlendedDiffuseSpecular    shading/closures/fresnelblendeddiffspec.cpp:8-69
//   SchlickDielectric / -Metallic    shading/fresnel/schlickdielectric.cpp, schlickmetallic.cpp
//   ThinFilmInteference              shading/fresnel/thinfilm.cpp:8-56, fresnel.h:9-23
//   EstevezSheenDistribution         shading/microfacets/estevez.cpp:8-42
//   ClearcoatReflectionBRDF          shading/bxdfs/clearcoat.cpp:7-81
//   BSDF evaluate / sample / pdf     shading/bsdf.cpp:12-151 (several lobes)
#pragma once
#include "fl_shade.h"

namespace fl {

enum : int { MAT_STANDARD = 4 };

enum : int {
    SL_OREN = 1,       // OrenNayarBRDF(kd, kdweight, kdroughness * 90)
    SL_DIFFTRANS = 2,  // DiffuseTransmission(kd, kdweight, kdroughness * 90) (thin walled)
    SL_FSPEC = 3,      // FresnelSpecularBXDF(ks, kt, 1, ksIOR)
    SL_CONDUCTOR = 4,  // MicrofacetReflection(GGX, FresnelConductor(forge metal), 1, ksweight)
    SL_FBLEND = 5,     // FresnelBlendedDiffuseSpecular(kd, kdweight, kdroughness, ks, ksweight, ro, Schlick, GGX)
    SL_THINFILM = 6,   // MicrofacetReflection(GGX, ThinFilmInteference(tfIOR, thickness, ksIOR), 1, 1)
    SL_SHEEN = 7,      // MicrofacetReflection(EstevezSheen(max(.001, sheenroughness)), FresnelNull, sheen, sheenweight)
    SL_COAT = 8        // ClearcoatReflectionBRDF(coat, coatIOR, coatweight, gloss)
};
constexpr int kStdMaxLobes = 5;

// Spectra a standard material's lobes read, per lane
enum : int { SS_KD = 0, SS_KS, SS_KT, SS_SHEEN, SS_COAT, SS_ETA, SS_K, SS_N };

// One material's lobe stack and constants (host-derived, as the reference's
// constructors derive them).
struct StdParams {
    int n;                      // lobes, in the order the reference adds them
    int lobe[kStdMaxLobes];
    float kd_weight, on_A, on_B;   // OrenNayar A, B for sigma = kdroughness * 90 (degrees)
    float fb_A, fb_B;              // the blended lobe's OrenNayar for sigma = kdroughness
    float ks_weight, ax, ay;       // GGX alphas (standardsurface.cpp:98-100)
    float ro;                      // ((1 - ksIOR) / (1 + ksIOR))^2
    float ks_ior;
    int fresnel_type;              // FL_FRESNEL_*
    float tf_ior, tf_thickness;
    float sheen_weight, sheen_rough;
    float coat_weight, coat_ior, coat_gloss;   // gloss after lerp(gloss, .1, .001)
    float eta_bsdf;
    uint32_t slot[SS_N];           // spectral table slots (render path)
    uint32_t tex[SS_ETA];          // 1 + image texture of kd, ks, kt, sheen, coat; 0 = the slot's constant
    uint32_t pad[3];
};

FL_HD uint32_t std_lobe_type(int l) {
    switch (l) {
    case SL_OREN: return BSDF_REFLECTION | BSDF_DIFFUSE;
    case SL_DIFFTRANS: return BSDF_TRANSMISSION | BSDF_DIFFUSE;
    case SL_FSPEC: return BSDF_REFLECTION | BSDF_TRANSMISSION | BSDF_SPECULAR;
    }
    return BSDF_REFLECTION | BSDF_GLOSSY;
}

FL_HD bool std_matches(int l, uint32_t flags) { return (std_lobe_type(l) & flags) == std_lobe_type(l); }

// OrenNayarBRDF::OrenNayarBRDF (orennayar.cpp:6-13): sigma in degrees
FL_HD void oren_nayar_AB(float sigma_deg, float *A, float *B) {
    const float sigma = sigma_deg * (kPi / 180.f);   // radians<float>
    const float sigma2 = sigma * sigma;
    *A = 1.f - (sigma2 / (2.f * (sigma2 + 0.33f)));
    *B = 0.45f * sigma2 / (sigma2 + 0.09f);
}

// OrenNayarBRDF::evaluate without R and weight: INV_PI * (A + B maxCos sinAlpha tanBeta)
FL_HD float oren_nayar_shape(float A, float B, v3 wo, v3 wi) {
    float sinThetaI = sint(wi), sinThetaO = sint(wo);
    float maxCos = 0.f;
    if (sinThetaI > 1e-4f && sinThetaO > 1e-4f) {
        float sinPhiI = sinphi(wi), cosPhiI = cosphi(wi);
        float sinPhiO = sinphi(wo), cosPhiO = cosphi(wo);
        float dCos = cosPhiI * cosPhiO + sinPhiI * sinPhiO;
        maxCos = fmaxf(0.f, dCos);
    }
    float sinAlpha, tanBeta;
    if (fabsf(wi.z) > fabsf(wo.z)) {
        sinAlpha = sinThetaO;
        tanBeta = sinThetaI / fabsf(wi.z);
    } else {
        sinAlpha = sinThetaI;
        tanBeta = sinThetaO / fabsf(wo.z);
    }
    return A + B * maxCos * sinAlpha * tanBeta;
}

// a . b with the error of each product (fma) and of each sum (TwoSum) carried
// along (the Dot2 algorithm of Ogita, Rump and Oishi, 2005): as accurate as a
// dot product in twice the precision. Used where a dot product of nearly
// orthogonal unit vectors divides the result.
FL_HD float dot_compensated(v3 a, v3 b) {
    float p = a.x * b.x;
    float e = fmaf(a.x, b.x, -p);
    float q = a.y * b.y;
    float s = p + q;
    float z = s - p;
    e += fmaf(a.y, b.y, -q) + ((p - (s - z)) + (q - z));
    float r = a.z * b.z;
    float t = s + r;
    z = t - s;
    e += fmaf(a.z, b.z, -r) + ((s - (t - z)) + (r - z));
    return t + e;
}

// ------------------------------------------------------------ Estevez ------
// sin(theta_h) as sqrt(x^2 + y^2) of the unit half vector: 1 - cos^2 cancels
// near the pole, and the power 1/r multiplies that relative error.
FL_HD float estevez_D(float r, v3 wh) {
    float invR = 1.f / r;
    return (2.f + invR) * powf(sqrtf(wh.x * wh.x + wh.y * wh.y), invR) / (2.f * kPi);
}

FL_HD float estevez_lambda(float r, v3 w) {
    const float q = (1.f - r) * (1.f - r);
    auto lerp_p = [q](float p0, float p1) { return q * p0 + (1.f - q) * p1; };
    const float a = lerp_p(25.3245f, 21.5473f), b = lerp_p(3.32435f, 3.82987f), c = lerp_p(0.16801f, 0.19823f),
                d = lerp_p(-1.27393f, -1.97760f), e = lerp_p(-4.85967f, -4.32054f);
    auto L = [a, b, c, d, e](float v) { return a / (1.f + b * powf(v, c)) + (d * v) + e; };
    float ac = fabsf(w.z);
    if (ac < 0.5f) return expf(L(ac));
    return expf(2.f * L(0.5f) - L(1.f - ac));
}

FL_HD float estevez_G(float r, v3 wo, v3 wi) { return 1.f / (1.f + estevez_lambda(r, wo) + estevez_lambda(r, wi)); }

// -------------------------------------------------------------- coat -------
// GTR1(|cos theta_h|, alpha) of clearcoat.cpp:8-12 for the unit half vector:
// 1 + (alpha^2 - 1) cos^2 = (x^2 + y^2) + alpha^2 z^2, which does not cancel
// near the pole (where a glossy coat puts its mass).
FL_HD float gtr1_wh(v3 wh, float alpha) {
    float alpha2 = alpha * alpha;
    return (alpha2 - 1.f) / (kPi * logf(alpha2) * ((wh.x * wh.x + wh.y * wh.y) + alpha2 * (wh.z * wh.z)));
}

FL_HD float smith_g_ggx(float cosTheta, float alpha) {
    float alpha2 = alpha * alpha;
    float cosTheta2 = cosTheta * cosTheta;
    return 1.f / (cosTheta + sqrtf(alpha2 + cosTheta2 - alpha2 * cosTheta2));
}

// ---------------------------------------------------------- thin film ------
// ThinFilmInteference::evaluate for one wavelength (nm), thinfilm.cpp:8-56
// with the amplitude coefficients of fresnel.h:9-23. The reference forms sin^2
// of both refracted angles as ((n0/n1)/(n0/n1)) (1 - cos^2), i.e. without the
// index ratio (thinfilm.cpp:10-11); x/x is exactly 1, so both refracted
// cosines are |cos theta_i| and the total-internal-reflection exit is never
// taken. That is kept.
//
// For cos theta_i = c > 0 every coefficient then loses its angle: with
// A = (n_f - 1)/(n_f + 1) and B = (n_m - n_f)/(n_m + n_f) both polarisations
// have alpha = A B, beta = (2/(1 + n_f)) (2 n_f/(n_f + n_m)) and the beam
// ratio is n_m, where beta^2 n_m = (1 - A^2)(1 - B^2). The reference's
// 1 - (ts + tp)/2 * beamRatio is therefore exactly
//     F = (A^2 + B^2 - 2AB cos phi') / (1 + A^2 B^2 - 2AB cos phi'),
// phi' = phi + delta. The device evaluates that form: 1 - T cancels to a few
// percent of T in fp32, this form does not (numerator and denominator are
// written as sums of non-negative terms, and delta in {0, pi, 2 pi} flips the
// sign of cos phi instead of being added to phi). c <= 0 (a back-side
// evaluation, and the reference's 0/0 at c = 0) keeps the reference's own
// arithmetic.
FL_HD float thin_film(float cosThetaI, float tfIOR, float thickness, float mediumIOR, float lambda_nm) {
    const float extIOR = 1.f;
    const bool flip = (tfIOR < extIOR) != (tfIOR < mediumIOR);   // delta = pi (mod 2 pi)
    if (cosThetaI > 0.f) {
        const float A = (tfIOR - extIOR) / (tfIOR + extIOR);
        const float B = (mediumIOR - tfIOR) / (mediumIOR + tfIOR);
        const float AB = A * B;
        float sh, ch;   // sin and cos of phi / 2 = (2 pi / lambda) n_f d c
        sincosf((2.f * kPi / lambda_nm) * (tfIOR * thickness * cosThetaI), &sh, &ch);
        const float s2 = sh * sh, c2 = ch * ch;
        const float one_minus = 2.f * (flip ? c2 : s2);   // 1 - cos(phi')
        const float one_plus = 2.f * (flip ? s2 : c2);    // 1 + cos(phi')
        // A^2 + B^2 - 2AB cos = (A - B)^2 + 2AB (1 - cos) = (A + B)^2 - 2AB (1 + cos)
        // 1 + (AB)^2 - 2AB cos = (1 - AB)^2 + 2AB (1 - cos) = (1 + AB)^2 - 2AB (1 + cos)
        const float num = AB >= 0.f ? (A - B) * (A - B) + 2.f * AB * one_minus : (A + B) * (A + B) - 2.f * AB * one_plus;
        const float den = AB >= 0.f ? (1.f - AB) * (1.f - AB) + 2.f * AB * one_minus
                                    : (1.f + AB) * (1.f + AB) - 2.f * AB * one_plus;
        return num / den;
    }
    float cos1 = fabsf(cosThetaI);
    float cos2 = cos1;
    auto sp_r = [](float n1, float n2, float cI, float cT) { return (n1 * cI - n2 * cT) / (n1 * cI + n2 * cT); };
    auto sp_t = [](float n1, float n2, float cI, float cT) { return (2.f * n1 * cI) / (n1 * cI + n2 * cT); };
    auto pp_r = [](float n1, float n2, float cI, float cT) { return (n2 * cI - n1 * cT) / (n1 * cT + n2 * cI); };
    auto pp_t = [](float n1, float n2, float cI, float cT) { return (2.f * n1 * cI) / (n1 * cT + n2 * cI); };
    float alphaSP = sp_r(tfIOR, extIOR, cos1, cosThetaI) * sp_r(mediumIOR, tfIOR, cos1, cos2);
    float alphaPP = pp_r(tfIOR, extIOR, cos1, cosThetaI) * pp_r(mediumIOR, tfIOR, cos1, cos2);
    float betaSP = sp_t(extIOR, tfIOR, cosThetaI, cos1) * sp_t(tfIOR, mediumIOR, cos1, cos2);
    float betaPP = pp_t(extIOR, tfIOR, cosThetaI, cos1) * pp_t(tfIOR, mediumIOR, cos1, cos2);
    float cphi = cosf((2.f * kPi / lambda_nm) * (2.f * tfIOR * thickness * cos1));
    if (flip) cphi = -cphi;
    float ts = (betaSP * betaSP) / ((alphaSP * alphaSP) - (2.f * alphaSP * cphi) + 1.f);
    float tp = (betaPP * betaPP) / ((alphaPP * alphaPP) - (2.f * alphaPP * cphi) + 1.f);
    float beamRatio = (mediumIOR * cos2) / (extIOR * cosThetaI);
    return 1.f - (ts + tp) * 0.5f * beamRatio;
}

// ------------------------------------------------------------- lobes -------
// Spec: float at(int which, int k) -- lane k of spectrum SS_*;
//       float tf_lambda(int k)   -- the wavelength (nm) the thin film sees on lane k.

template <class Spec>
FL_HD Bsdf std_conductor(const StdParams &m, const Spec &sp, bool vndf) {
    Bsdf b;
    b.kind = MAT_CONDUCTOR;
    b.ggx.ax = m.ax;
    b.ggx.ay = m.ay;
    b.weight = m.ks_weight;
    b.vndf_pdf = vndf;
    for (int k = 0; k < 4; ++k) { b.s1[k] = sp.at(SS_ETA, k); b.s2[k] = sp.at(SS_K, k); }
    b.ior = 1.f;
    b.dispersive = false;
    return b;
}

template <class Spec>
FL_HD Bsdf std_fspec(const StdParams &m, const Spec &sp) {
    Bsdf b;
    b.kind = MAT_DIELECTRIC;
    b.ggx.ax = b.ggx.ay = 1.f;
    b.weight = 1.f;
    b.vndf_pdf = false;
    for (int k = 0; k < 4; ++k) { b.s1[k] = sp.at(SS_KS, k); b.s2[k] = sp.at(SS_KT, k); }
    b.ior = m.ks_ior;
    b.dispersive = false;
    return b;
}

// BxDF::evaluate of lobe l (local frame), f per lane
template <class Spec>
FL_HD void std_lobe_eval(const StdParams &m, int l, const Spec &sp, bool vndf, v3 wo, v3 wi, float f[4]) {
    for (int k = 0; k < 4; ++k) f[k] = 0.f;
    switch (l) {
    case SL_OREN:
    case SL_DIFFTRANS: {
        const float s = oren_nayar_shape(m.on_A, m.on_B, wo, wi);
        for (int k = 0; k < 4; ++k) f[k] = sp.at(SS_KD, k) * kInvPi * s * m.kd_weight;
    } break;
    case SL_CONDUCTOR: {
        Bsdf b = std_conductor(m, sp, vndf);
        microfacet_f(b, wo, wi, f);
    } break;
    case SL_FBLEND: {
        // 1 - (1 - a)^5 with a = |cos| / 2 (fresnelblendeddiffspec.cpp:22-27)
        // as a (1 + r + r^2 + r^3 + r^4), r = 1 - a: near grazing the
        // difference of 1 and r^5 keeps only a few digits in fp32
        auto one_minus_pow5 = [](float c) {
            const float a = .5f * fabsf(c), r = 1.f - a;
            return a * (1.f + r * (1.f + r * (1.f + r * (1.f + r))));
        };
        const float d1 = one_minus_pow5(wi.z), d2 = one_minus_pow5(wo.z);
        v3 wh = wi + wo;
        if (wh.x == 0.f && wh.y == 0.f && wh.z == 0.f) return;
        wh = normalize(wh);
        const float on = oren_nayar_shape(m.fb_A, m.fb_B, wo, wi);
        GGX g{m.ax, m.ay};
        const float D = ggx_D(g, wh);
        // near grazing retro-reflection wi . wh is small and divides the
        // specular term: plain fp32 would cost it ~1e-5 of its precision
        const float cwh = dot_compensated(wi, wh);
        const float den = 4.f * fabsf(cwh) * fmaxf(fabsf(wi.z), fabsf(wo.z));
        float R1 = 1.f - cwh;
        const float p5 = (R1 * R1) * (R1 * R1) * R1;                 // SchlickDielectric
        const float Fd = m.ro + p5 * (1.f - m.ro);
        const float p5m = R1 * R1 * R1 * R1 * R1;                     // SchlickMetallic
        for (int k = 0; k < 4; ++k) {
            const float ks = sp.at(SS_KS, k);
            const float diffuse =
                (28.f / 23.f) * (sp.at(SS_KD, k) * kInvPi * on * m.kd_weight) * (1.f - m.ro) * d1 * d2;
            const float F = m.fresnel_type == 1 /* FL_FRESNEL_METALLIC */ ? ks + (1.f - ks) * p5m : Fd;
            float spec = D * F / den;
            spec = ks * spec * m.ks_weight;
            f[k] = diffuse + spec;
        }
    } break;
    case SL_THINFILM:
    case SL_SHEEN: {
        // MicrofacetReflectionBRDF::evaluate (microfacetreflection.cpp:13-32)
        float cO = fabsf(wo.z), cI = fabsf(wi.z);
        v3 wh = wi + wo;
        if (cI == 0.f || cO == 0.f) return;
        if (wh.x == 0.f && wh.y == 0.f && wh.z == 0.f) return;
        wh = normalize(wh);
        if (l == SL_THINFILM) {
            GGX g{m.ax, m.ay};
            const float D = ggx_D(g, wh), G = ggx_G(g, wo, wi);
            const float cosF = dot(wi, face_forward(wh, v3(0.f, 0.f, 1.f)));
            for (int k = 0; k < 4; ++k) {
                const float F = thin_film(cosF, m.tf_ior, m.tf_thickness, m.ks_ior, sp.tf_lambda(k));
                f[k] = 1.f * 1.f * D * G * F / (4.f * cO * cI);
            }
        } else {
            const float D = estevez_D(m.sheen_rough, wh), G = estevez_G(m.sheen_rough, wo, wi);
            for (int k = 0; k < 4; ++k) f[k] = sp.at(SS_SHEEN, k) * m.sheen_weight * D * G * 1.f / (4.f * cO * cI);
        }
    } break;
    case SL_COAT: {
        v3 wh = wi + wo;
        if (wh.x == 0.f && wh.y == 0.f && wh.z == 0.f) return;
        wh = normalize(wh);
        const float Dr = gtr1_wh(wh, m.coat_gloss);
        const float Fr = fresnel_dielectric(dot(wo, wh), 1.f, m.coat_ior);
        const float Gr = smith_g_ggx(fabsf(wo.z), .25f) * smith_g_ggx(fabsf(wi.z), .25f);
        for (int k = 0; k < 4; ++k) f[k] = sp.at(SS_COAT, k) * Fr * m.coat_weight * Gr * Dr / 4.f;
    } break;
    default: break;   // SL_FSPEC: evaluate is 0
    }
}

// BxDF::pdf of lobe l
FL_HD float std_lobe_pdf(const StdParams &m, int l, bool vndf, v3 wo, v3 wi) {
    switch (l) {
    case SL_OREN: return same_hemi(wo, wi) ? fabsf(wi.z) * kInvPi : 0.f;
    case SL_DIFFTRANS: return !same_hemi(wo, wi) ? fabsf(wi.z) * kInvPi : 0.f;
    case SL_CONDUCTOR:
    case SL_THINFILM: {
        if (!same_hemi(wo, wi)) return 0.f;
        v3 wh = normalize(wo + wi);
        GGX g{m.ax, m.ay};
        return ggx_pdf_wh(g, wo, wh, vndf) / (4.f * dot(wo, wh));
    }
    case SL_FBLEND: {
        if (!same_hemi(wo, wi)) return 0.f;
        v3 wh = normalize(wo + wi);
        GGX g{m.ax, m.ay};
        float pdf_wh = ggx_pdf_wh(g, wo, wh, vndf);
        return .5f * (fabsf(wi.z) * kInvPi + pdf_wh / (4.f * dot(wo, wh)));
    }
    case SL_SHEEN: {
        if (!same_hemi(wo, wi)) return 0.f;
        v3 wh = normalize(wo + wi);
        return estevez_D(m.sheen_rough, wh) * fabsf(wh.z) / (4.f * dot(wo, wh));   // MicrofacetDistribution::pdf
    }
    case SL_COAT: {
        if (!same_hemi(wo, wi)) return 0.f;
        v3 wh = wi + wo;
        if (wh.x == 0.f && wh.y == 0.f && wh.z == 0.f) return 0.f;
        wh = normalize(wh);
        float Dr = gtr1_wh(wh, m.coat_gloss);
        return Dr * fabsf(wh.z) / (4.f * dot(wo, wh));
    }
    }
    return 0.f;
}

// BxDF::sample of lobe l; s.pdf stays 0 on failure
template <class Spec>
FL_HD void std_lobe_sample(const StdParams &m, int l, const Spec &sp, bool vndf, v3 wo, float u0, float u1,
                           BsdfSample &s) {
    s.pdf = 0.f;
    s.type = std_lobe_type(l);
    s.secondary_terminated = false;
    for (int k = 0; k < 4; ++k) s.f[k] = 0.f;
    switch (l) {
    case SL_OREN:
    case SL_DIFFTRANS: {
        v3 wi = cosine_hemisphere(u0, u1);
        if (l == SL_OREN ? wo.z < 0.f : wo.z > 0.f) wi.z *= -1.f;
        s.wi = wi;
        s.pdf = std_lobe_pdf(m, l, vndf, wo, wi);
        std_lobe_eval(m, l, sp, vndf, wo, wi, s.f);
    } break;
    case SL_CONDUCTOR: {
        Bsdf b = std_conductor(m, sp, vndf);
        lobe_sample(b, wo, u0, u1, s);
    } break;
    case SL_FSPEC: {
        Bsdf b = std_fspec(m, sp);
        lobe_sample(b, wo, u0, u1, s);
    } break;
    case SL_FBLEND: {
        v3 wi;
        if (u0 < .5f) {
            float ua = fminf(2.f * u0, 1.f - kShadowEps);
            wi = cosine_hemisphere(ua, u1);
            if (wo.z < 0.f) wi.z *= -1.f;
        } else {
            float ua = fminf(2.f * (u0 - .5f), 1.f - kShadowEps);
            GGX g{m.ax, m.ay};
            v3 wh = ggx_sample_wh(g, wo, ua, u1);
            wi = reflect(wo, wh);
            if (!same_hemi(wo, wi)) return;
        }
        s.wi = wi;
        s.pdf = std_lobe_pdf(m, l, vndf, wo, wi);
        std_lobe_eval(m, l, sp, vndf, wo, wi, s.f);
    } break;
    case SL_THINFILM:
    case SL_SHEEN: {
        if (wo.z == 0.f) return;
        v3 wh;
        if (l == SL_THINFILM) {
            GGX g{m.ax, m.ay};
            wh = ggx_sample_wh(g, wo, u0, u1);
        } else {
            wh = cosine_hemisphere(u0, u1);          // EstevezSheenDistribution::sample_wh
            if (wo.z < 0.f) wh.z *= -1.f;
        }
        if (dot(wo, wh) < 0.f) return;
        v3 wi = reflect(wo, wh);
        if (!same_hemi(wo, wi)) return;
        s.wi = wi;
        if (l == SL_THINFILM) {
            GGX g{m.ax, m.ay};
            s.pdf = ggx_pdf_wh(g, wo, wh, vndf) / (4.f * dot(wo, wh));
        } else {
            s.pdf = estevez_D(m.sheen_rough, wh) * fabsf(wh.z) / (4.f * dot(wo, wh));
        }
        std_lobe_eval(m, l, sp, vndf, wo, wi, s.f);
    } break;
    case SL_COAT: {
        if (wo.z == 0.f) return;
        // clearcoat.cpp:47-56: cos^2 = (1 - a^(2(1-u))) / (1 - a^2). With
        // a^(2(1-u)) = exp(2 (1-u) ln a), cos^2 and sin^2 = 1 - cos^2 are both
        // expm1 forms, which keep their precision at either end of u.
        float alpha2 = m.coat_gloss * m.coat_gloss;
        float lnA = logf(m.coat_gloss);
        float cos2 = -expm1f(2.f * (1.f - u0) * lnA) / (1.f - alpha2);
        float sin2 = alpha2 * expm1f(-2.f * u0 * lnA) / (1.f - alpha2);
        float cosTheta = sqrtf(fmaxf(0.f, cos2));
        float sinTheta = sqrtf(fmaxf(0.f, sin2));
        float phi = 2.f * kPi * u1;
        v3 wh(sinTheta * cosf(phi), sinTheta * sinf(phi), cosTheta);
        if (!same_hemi(wo, wh)) wh = -wh;
        v3 wi = reflect(wo, wh);
        if (!same_hemi(wo, wi)) return;
        s.wi = wi;
        s.pdf = std_lobe_pdf(m, l, vndf, wo, wi);
        std_lobe_eval(m, l, sp, vndf, wo, wi, s.f);
    } break;
    }
}

// ------------------------------------------------------------- BSDF --------
FL_HD int std_num_components(const StdParams &m, uint32_t flags) {
    int n = 0;
    for (int i = 0; i < m.n; ++i) n += std_matches(m.lobe[i], flags) ? 1 : 0;
    return n;
}

// BSDF::evaluate(wo, wi, flags) in the local frame (ns = ng, so reflect is
// wo.z * wi.z > 0; the caller passes the world-space test)
template <class Spec>
FL_HD void std_bsdf_eval(const StdParams &m, const Spec &sp, bool vndf, v3 wo, v3 wi, bool reflect, uint32_t flags,
                         float f[4]) {
    for (int k = 0; k < 4; ++k) f[k] = 0.f;
    if (wo.z == 0.f) return;
    for (int i = 0; i < m.n; ++i) {
        const int l = m.lobe[i];
        const uint32_t t = std_lobe_type(l);
        if (std_matches(l, flags) && ((reflect && (t & BSDF_REFLECTION)) || (!reflect && (t & BSDF_TRANSMISSION)))) {
            float g[4];
            std_lobe_eval(m, l, sp, vndf, wo, wi, g);
            for (int k = 0; k < 4; ++k) f[k] += g[k];
        }
    }
}

// BSDF::pdf(wo, wi, flags)
FL_HD float std_bsdf_pdf(const StdParams &m, bool vndf, v3 wo, v3 wi, uint32_t flags) {
    if (m.n == 0 || wo.z == 0.f) return 0.f;
    float pdf = 0.f;
    int matching = 0;
    for (int i = 0; i < m.n; ++i)
        if (std_matches(m.lobe[i], flags)) {
            ++matching;
            pdf += std_lobe_pdf(m, m.lobe[i], vndf, wo, wi);
        }
    return matching > 0 ? pdf / (float)matching : 0.f;
}

// BSDF::sample(wo, u, flags) in the local frame (bsdf.cpp:54-126): s.pdf = 0
// on failure; s.type the sampled lobe's type.
template <class Spec>
FL_HD void std_bsdf_sample(const StdParams &m, const Spec &sp, bool vndf, v3 wo, float u0, float u1, uint32_t flags,
                           BsdfSample &s) {
    s.pdf = 0.f;
    s.type = 0;
    s.secondary_terminated = false;
    for (int k = 0; k < 4; ++k) s.f[k] = 0.f;
    const int matchComps = std_num_components(m, flags);
    if (matchComps == 0) return;
    const int c0 = (int)floorf(u0 * (float)matchComps);
    const int comp = c0 < matchComps - 1 ? c0 : matchComps - 1;
    int chosen = -1, count = comp;
    for (int i = 0; i < m.n; ++i)
        if (std_matches(m.lobe[i], flags) && count-- == 0) {
            chosen = i;
            break;
        }
    const float ua = fminf(u0 * (float)matchComps - (float)comp, 1.f - kShadowEps);
    if (wo.z == 0.f) return;
    const int l = m.lobe[chosen];
    std_lobe_sample(m, l, sp, vndf, wo, ua, u1, s);
    if (s.pdf == 0.f) {
        s.type = 0;
        return;
    }
    const v3 wi = s.wi;
    const bool specular = (std_lobe_type(l) & BSDF_SPECULAR) != 0;
    if (!specular && matchComps > 1)
        for (int i = 0; i < m.n; ++i)
            if (i != chosen && std_matches(m.lobe[i], flags)) s.pdf += std_lobe_pdf(m, m.lobe[i], vndf, wo, wi);
    if (matchComps > 1) s.pdf /= (float)matchComps;
    if (!specular) std_bsdf_eval(m, sp, vndf, wo, wi, wo.z * wi.z > 0.f, flags, s.f);
}

} // namespace fl
