// fl_guide.h -- the learned directional proposal: network shape, input
// encoding, and the von Mises-Fisher mixture it predicts.
//
// Network: 53 inputs (position in the scene box, 8 frequency bands per axis,
// sin and cos: 48; shading normal: 3; wavelength: 1; roughness: 1), padded to
// 64, then 4 hidden layers of 64 ReLU units, then 40 outputs = 8 lobes x
// (mean direction 3, concentration 1, weight logit 1). Weights are bf16 on the
// matrix cores with fp32 accumulation and fp32 master copies. All five weight
// matrices are stored 64 x 64 (out x in): the unused input columns see zero
// inputs and the unused output rows are never read, so both stay inert.
//
// vMF lobe:  v(w) = kappa / (2 pi (1 - exp(-2 kappa))) * exp(kappa (mu.w - 1))
// (normalized on the sphere; written in the form that does not overflow).
// Sampling: W. Jakob, "Numerically stable sampling of the von Mises-Fisher
// distribution on S^2" (2012): w = 1 + log(u + (1 - u) exp(-2 kappa)) / kappa.
#pragma once
#include "fl_math.h"

namespace fl {

constexpr int kGH = 64;                              // layer width (and padded input width)
constexpr int kGMats = 5;                            // weight matrices: 4 hidden layers + head
constexpr int kGIn = 53;                             // inputs actually used
constexpr int kGLobes = 8;
constexpr int kGOut = kGLobes * 5;                   // 40 raw outputs
constexpr int kGBands = 8;
constexpr int kGLayerParams = kGH * kGH + kGH;       // W (64x64, out-major) then b (64)
constexpr int kGParams = kGMats * kGLayerParams;     // 20,800 stored (18,536 algorithmic)
constexpr float kGKappaLo = -3.f, kGKappaHi = 7.f;   // kappa = exp(clamp(raw, lo, hi))

struct VmfLobe {
    v3 mu;
    float kappa;
    float w;        // mixture weight (softmax), sums to 1 over lobes
};

// Raw head outputs (40 floats, lobe-major: mu xyz, kappa raw, weight logit)
// -> normalized lobes.
FL_HD void guide_decode(const float *raw, VmfLobe *lobes) {
    float mx = -3.0e38f;
    for (int i = 0; i < kGLobes; ++i) mx = fmaxf(mx, raw[i * 5 + 4]);
    float sum = 0.f;
    for (int i = 0; i < kGLobes; ++i) {
        const float *r = raw + i * 5;
        float len = sqrtf(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
        if (len > 1e-6f) lobes[i].mu = v3(r[0] / len, r[1] / len, r[2] / len);
        else lobes[i].mu = v3(0.f, 0.f, 1.f);
        lobes[i].kappa = expf(fminf(fmaxf(r[3], kGKappaLo), kGKappaHi));
        lobes[i].w = expf(r[4] - mx);
        sum += lobes[i].w;
    }
    for (int i = 0; i < kGLobes; ++i) lobes[i].w = lobes[i].w / sum;
}

// Normalization constant kappa / (2 pi (1 - exp(-2 kappa))); the limit
// 1 / (4 pi) at kappa -> 0 comes out of expm1 without cancellation.
FL_HD float vmf_norm(float kappa) { return kappa / (2.f * kPi * -expm1f(-2.f * kappa)); }

FL_HD float vmf_pdf(const VmfLobe &l, v3 w) { return vmf_norm(l.kappa) * expf(l.kappa * (dot(l.mu, w) - 1.f)); }

FL_HD float vmf_mixture_pdf(const VmfLobe *lobes, v3 w) {
    float p = 0.f;
    for (int i = 0; i < kGLobes; ++i) p += lobes[i].w * vmf_pdf(lobes[i], w);
    return p;
}

// Direction from lobe `l` for uniforms (u1, u2) in [0,1).
FL_HD v3 vmf_sample(const VmfLobe &l, float u1, float u2) {
    float wz;
    if (l.kappa < 1e-4f) wz = 1.f - 2.f * u1;   // uniform sphere limit
    else wz = 1.f + logf(u1 + (1.f - u1) * expf(-2.f * l.kappa)) / l.kappa;
    wz = fminf(fmaxf(wz, -1.f), 1.f);
    const float r = sqrtf(fmaxf(0.f, 1.f - wz * wz));
    const float phi = 2.f * kPi * u2;
    v3 s, t;
    coordinate_system(l.mu, &s, &t);
    s = normalize(s);
    t = normalize(t);
    return normalize(s * (r * cosf(phi)) + t * (r * sinf(phi)) + l.mu * wz);
}

// Picks a lobe by its weight with u in [0,1) and returns u rescaled to [0,1)
// within the chosen lobe's interval.
FL_HD int vmf_pick(const VmfLobe *lobes, float u, float *u_re) {
    float c = 0.f;
    for (int i = 0; i < kGLobes - 1; ++i) {
        const float nc = c + lobes[i].w;
        if (u < nc) {
            *u_re = fminf((u - c) / fmaxf(lobes[i].w, 1e-30f), 0x1.fffffep-1f);
            return i;
        }
        c = nc;
    }
    *u_re = fminf((u - c) / fmaxf(lobes[kGLobes - 1].w, 1e-30f), 0x1.fffffep-1f);
    *u_re = fmaxf(*u_re, 0.f);
    return kGLobes - 1;
}

} // namespace fl
