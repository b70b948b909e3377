// oracle_guide.cpp -- CPU oracle of the learned proposal (TEST INFRASTRUCTURE).
//
// FP64 restatement of the guiding network the GPU trains: 64 encoded inputs,
// 4 hidden ReLU layers of 64, a 40-value head decoded into 8 vMF lobes, the
// loss -c log p(dir) and its gradient by backpropagation. With bf16 = 1 the
// values the GPU rounds to bf16 before its matrix products (inputs, weights,
// hidden activations, output deltas) are rounded here at the same points, so
// the GPU gradient differs from this one only by fp32 accumulation order.
// With bf16 = 0 everything is exact FP64, which the tests compare with
// central finite differences of the loss.
//
// The reference has no learned proposal; this restates the build's own
// definition (vMF normalisation and sampling after W. Jakob, "Numerically
// stable sampling of the von Mises-Fisher distribution on S^2", 2012).
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "oracle.h"

namespace {

constexpr int H = 64, NM = 5, NOUT = 40, NL = 8, LP = H * H + H;
constexpr double KLO = -3.0, KHI = 7.0;

double bf(double x, int on) {
    if (!on) return x;
    float f = (float)x;
    uint32_t u;
    std::memcpy(&u, &f, 4);
    if ((u & 0x7f800000u) != 0x7f800000u) {   // finite: round to nearest even on 16 bits
        u += 0x7fffu + ((u >> 16) & 1u);
        u &= 0xffff0000u;
    }
    std::memcpy(&f, &u, 4);
    return f;
}

// fp32 value of an accumulated sum (the GPU accumulates in fp32)
double f32(double x, int on) { return on ? (double)(float)x : x; }

struct Lobe {
    double mu[3], kappa, w;
};

void decode(const double *raw, Lobe *L) {
    double mx = -1e300;
    for (int i = 0; i < NL; ++i) mx = std::fmax(mx, raw[5 * i + 4]);
    double s = 0.0;
    for (int i = 0; i < NL; ++i) {
        const double *r = raw + 5 * i;
        double len = std::sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
        if (len > 1e-6) for (int k = 0; k < 3; ++k) L[i].mu[k] = r[k] / len;
        else { L[i].mu[0] = 0; L[i].mu[1] = 0; L[i].mu[2] = 1; }
        L[i].kappa = std::exp(std::fmin(std::fmax(r[3], KLO), KHI));
        L[i].w = std::exp(r[4] - mx);
        s += L[i].w;
    }
    for (int i = 0; i < NL; ++i) L[i].w /= s;
}

double vmf(const Lobe &l, const double *d) {
    const double md = l.mu[0] * d[0] + l.mu[1] * d[1] + l.mu[2] * d[2];
    return l.kappa / (2.0 * M_PI * -std::expm1(-2.0 * l.kappa)) * std::exp(l.kappa * (md - 1.0));
}

// loss -c log p and d loss / d raw
double head(const double *raw, const double *d, double c, double *g) {
    Lobe L[NL];
    decode(raw, L);
    double v[NL], p = 0.0;
    for (int i = 0; i < NL; ++i) { v[i] = vmf(L[i], d); p += L[i].w * v[i]; }
    p = std::fmax(p, 1e-300);
    if (g) {
        for (int i = 0; i < NL; ++i) {
            const double *r = raw + 5 * i;
            double *gi = g + 5 * i;
            const double resp = L[i].w * v[i] / p;
            gi[4] = -c * (resp - L[i].w);
            const double k = L[i].kappa;
            const double md = L[i].mu[0] * d[0] + L[i].mu[1] * d[1] + L[i].mu[2] * d[2];
            const bool inside = r[3] > KLO && r[3] < KHI;
            gi[3] = inside ? -c * resp * (1.0 / k - 2.0 / std::expm1(2.0 * k) + (md - 1.0)) * k : 0.0;
            const double len = std::sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
            for (int a = 0; a < 3; ++a) gi[a] = len > 1e-6 ? -c * resp * k * (d[a] - md * L[i].mu[a]) / len : 0.0;
        }
    }
    return -c * std::log(p);
}

// forward of one sample; A[k] = input of matrix k (A[0] = x), raw = head
void forward(const double *P, const float *x, int b16, double A[NM][H], double *raw) {
    for (int i = 0; i < H; ++i) A[0][i] = bf(x[i], b16);
    for (int k = 0; k < NM; ++k) {
        const double *W = P + k * LP, *b = W + H * H;
        for (int o = 0; o < H; ++o) {
            double s = 0.0;
            for (int i = 0; i < H; ++i) s += bf(W[o * H + i], b16) * A[k][i];
            double z = f32(s, b16) + b[o];
            z = f32(z, b16);
            if (k < NM - 1) A[k + 1][o] = bf(std::fmax(z, 0.0), b16);
            else raw[o] = z;
        }
    }
}

} // namespace

extern "C" {

double or_bf16_round(double x) { return bf(x, 1); }

void or_guide_encode(const float pos01[3], const float n[3], float lam01, float rough, float *x) {
    for (int a = 0; a < 3; ++a)
        for (int k = 0; k < 8; ++k) {
            const double f = std::ldexp((double)pos01[a], k) * M_PI;
            x[16 * a + 2 * k] = (float)std::sin(f);
            x[16 * a + 2 * k + 1] = (float)std::cos(f);
        }
    x[48] = n[0]; x[49] = n[1]; x[50] = n[2];
    x[51] = lam01;
    x[52] = rough;
    for (int i = 53; i < 64; ++i) x[i] = 0.f;
}

void or_guide_forward(const double *P, const float *x, uint32_t n, int b16, double *raw) {
    double A[NM][H], r[H];
    for (uint32_t s = 0; s < n; ++s) {
        forward(P, x + (size_t)s * 64, b16, A, r);
        for (int o = 0; o < NOUT; ++o) raw[(size_t)s * NOUT + o] = r[o];
    }
}

// records: n x 68 (x[64], dir[3], c); returns the loss sum, grad (if given)
// receives the gradient sum (kGParams = 5 x 4160 doubles)
double or_guide_gradient(const double *P, const float *rec, uint32_t n, int b16, double *grad) {
    if (grad) std::memset(grad, 0, sizeof(double) * NM * LP);
    double loss = 0.0;
    double A[NM][H], raw[H], g[NOUT];
    double delta[H], dA[H];
    for (uint32_t s = 0; s < n; ++s) {
        const float *r = rec + (size_t)s * 68;
        forward(P, r, b16, A, raw);
        const double d[3] = {r[64], r[65], r[66]};
        const double c = r[67];
        if (!(c > 0.0)) continue;
        loss += head(raw, d, c, grad ? g : nullptr);
        if (!grad) continue;
        for (int o = 0; o < H; ++o) delta[o] = o < NOUT ? f32(g[o], b16) : 0.0;
        for (int k = NM - 1; k >= 0; --k) {
            double *G = grad + k * LP;
            const double *W = P + k * LP;
            for (int o = 0; o < H; ++o) {
                G[H * H + o] += delta[o];                      // bias: unrounded delta
                const double db = bf(delta[o], b16);
                for (int i = 0; i < H; ++i) G[o * H + i] += db * A[k][i];
            }
            if (k == 0) break;
            for (int i = 0; i < H; ++i) {
                double s2 = 0.0;
                for (int o = 0; o < H; ++o) s2 += bf(delta[o], b16) * bf(W[o * H + i], b16);
                dA[i] = A[k][i] > 0.0 ? f32(s2, b16) : 0.0;   // ReLU mask of A_k = relu(Z_{k-1})
            }
            for (int i = 0; i < H; ++i) delta[i] = dA[i];
        }
    }
    return loss;
}

// mixture pdf of decoded raw head values at a unit direction
double or_vmf_mixture_pdf(const double *raw, const double *dir) {
    Lobe L[NL];
    decode(raw, L);
    double p = 0.0;
    for (int i = 0; i < NL; ++i) p += L[i].w * vmf(L[i], dir);
    return p;
}

// one lobe's normalisation constant and a sample from it (local frame: mu = z)
double or_vmf_norm(double kappa) { return kappa / (2.0 * M_PI * -std::expm1(-2.0 * kappa)); }

double or_vmf_sample_cos(double kappa, double u) {
    if (kappa < 1e-4) return 1.0 - 2.0 * u;
    return 1.0 + std::log(u + (1.0 - u) * std::exp(-2.0 * kappa)) / kappa;
}

} // extern "C"
