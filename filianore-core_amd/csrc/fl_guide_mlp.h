// fl_guide_mlp.h -- matrix-core building blocks of the guiding network on
// gfx950: one wave owns 64 samples; every product is 64 x 64 x 64 and runs
// as 4 x 4 output tiles of v_mfma_f32_16x16x32_bf16 over 2 K steps.
//
// Operand layout of v_mfma_f32_16x16x32_bf16 (D = A B, A 16x32, B 32x16):
//   lane l supplies A[l % 16][8 (l / 16) + j] and B[8 (l / 16) + j][l % 16],
//   j = 0..7 (one 16-byte bf16x8 each), and receives D[4 (l / 16) + r][l % 16]
//   in result register r = 0..3.
// Every product here is written as Out^T = P Q^T with P stored [m][k] and Q
// stored [n][k] (both k-contiguous), so each fragment is one 16-byte load and
// each lane's 4 results are 4 consecutive m for one n.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "fl_guide.h"

namespace fl {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int kGRow = 72;   // bf16 row stride of 64-wide tiles in LDS (144 B: spreads rows over banks)

__device__ __forceinline__ __bf16 to_bf16(float x) { return (__bf16)x; }   // round to nearest even

// acc[mt][nt] (+)= P[m][k] * Q[n][k] summed over k = 0..63, for m, n in 0..63.
// P row stride ps, Q row stride qs (in bf16 elements); both may be LDS or global.
__device__ __forceinline__ void mm64(const __bf16 *P, int ps, const __bf16 *Q, int qs, f32x4 acc[4][4]) {
    const int l = threadIdx.x & 63;
    const int r16 = l & 15, kq = (l >> 4) * 8;
#pragma unroll
    for (int ks = 0; ks < 2; ++ks) {
        bf16x8 a[4], b[4];
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            a[t] = *reinterpret_cast<const bf16x8 *>(P + (t * 16 + r16) * ps + ks * 32 + kq);
            b[t] = *reinterpret_cast<const bf16x8 *>(Q + (t * 16 + r16) * qs + ks * 32 + kq);
        }
#pragma unroll
        for (int mt = 0; mt < 4; ++mt)
#pragma unroll
            for (int nt = 0; nt < 4; ++nt)
                acc[mt][nt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[mt], b[nt], acc[mt][nt], 0, 0, 0);
    }
}

// acc[mt][nt] (+)= sum_k P[16 mt + i][k] * Q[16 nt + j][k] over k = 0..63:
// MT x NT output tiles of 16 x 16 (rows of P, rows of Q), P and Q already
// pointing at the first row of the tiles. Element r of acc[mt][nt] is row
// 16 mt + 4 (lane / 16) + r, column 16 nt + lane % 16.
template <int MT, int NT>
__device__ __forceinline__ void mm_tiles(const __bf16 *P, int ps, const __bf16 *Q, int qs, f32x4 (&acc)[MT][NT]) {
    const int l = threadIdx.x & 63;
    const int r16 = l & 15, kq = (l >> 4) * 8;
#pragma unroll
    for (int ks = 0; ks < 2; ++ks) {
        bf16x8 a[MT], b[NT];
#pragma unroll
        for (int t = 0; t < MT; ++t) a[t] = *reinterpret_cast<const bf16x8 *>(P + (t * 16 + r16) * ps + ks * 32 + kq);
#pragma unroll
        for (int t = 0; t < NT; ++t) b[t] = *reinterpret_cast<const bf16x8 *>(Q + (t * 16 + r16) * qs + ks * 32 + kq);
#pragma unroll
        for (int mt = 0; mt < MT; ++mt)
#pragma unroll
            for (int nt = 0; nt < NT; ++nt)
                acc[mt][nt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[mt], b[nt], acc[mt][nt], 0, 0, 0);
    }
}

template <int MT, int NT>
__device__ __forceinline__ void acc_zero_t(f32x4 (&acc)[MT][NT]) {
#pragma unroll
    for (int mt = 0; mt < MT; ++mt)
#pragma unroll
        for (int nt = 0; nt < NT; ++nt) acc[mt][nt] = f32x4{0.f, 0.f, 0.f, 0.f};
}

// the 16 encoded inputs [16 j, 16 j + 16) of a sample (j = 0..3): axis j's 8
// bands of sin and cos for j < 3 (guide_encode's double-angle recurrence),
// then normal, wavelength, roughness and zeros for j = 3
__device__ __forceinline__ void guide_encode_quarter(int j, const float pos01[3], v3 n, float lam01, float rough,
                                                     float *x16) {
    if (j < 3) {
        float sn, cs;
        sincospif(pos01[j], &sn, &cs);
#pragma unroll
        for (int k = 0; k < kGBands; ++k) {
            x16[2 * k] = sn;
            x16[2 * k + 1] = cs;
            const float s2 = 2.f * sn * cs;
            cs = (cs - sn) * (cs + sn);
            sn = s2;
        }
    } else {
        x16[0] = n.x;
        x16[1] = n.y;
        x16[2] = n.z;
        x16[3] = lam01;
        x16[4] = rough;
#pragma unroll
        for (int i = 5; i < 16; ++i) x16[i] = 0.f;
    }
}

__device__ __forceinline__ void acc_zero(f32x4 acc[4][4]) {
#pragma unroll
    for (int mt = 0; mt < 4; ++mt)
#pragma unroll
        for (int nt = 0; nt < 4; ++nt) acc[mt][nt] = f32x4{0.f, 0.f, 0.f, 0.f};
}

// Row m of acc element (mt, nt, r) for this lane, and its column n.
__device__ __forceinline__ int acc_m(int mt, int r) { return mt * 16 + ((threadIdx.x & 63) >> 4) * 4 + r; }
__device__ __forceinline__ int acc_n(int nt) { return nt * 16 + (threadIdx.x & 15); }

// Input encoding of one sample (53 used values, zero-padded to 64):
//   [16 a + 2 k] = sin(2^k pi x_a), [16 a + 2 k + 1] = cos(2^k pi x_a)  (a = 0..2, k = 0..7)
//   [48..50] = normal, [51] = wavelength in [0,1], [52] = roughness, rest 0.
// Band k + 1 comes from band k by the double-angle identities
// sin 2t = 2 sin t cos t, cos 2t = (cos t - sin t)(cos t + sin t): one sincos
// per axis instead of eight. Each doubling at most doubles the absolute error,
// so band 7 is within ~1e-5 of sin/cos of 2^7 pi x; the network sees bf16
// inputs (spacing 2^-8 near 1), so the difference does not reach it.
__device__ __forceinline__ void guide_encode(const float pos01[3], v3 n, float lam01, float rough, float *x) {
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        float sn, cs;
        sincospif(pos01[a], &sn, &cs);
#pragma unroll
        for (int k = 0; k < kGBands; ++k) {
            x[16 * a + 2 * k] = sn;
            x[16 * a + 2 * k + 1] = cs;
            const float s2 = 2.f * sn * cs;
            cs = (cs - sn) * (cs + sn);
            sn = s2;
        }
    }
    x[48] = n.x;
    x[49] = n.y;
    x[50] = n.z;
    x[51] = lam01;
    x[52] = rough;
#pragma unroll
    for (int i = 53; i < 64; ++i) x[i] = 0.f;
}

// Stores a sample's 64 encoded inputs as bf16 into row q of a [64][kGRow]
// LDS tile (eight 16-byte stores; kGRow * 2 bytes is a multiple of 16).
__device__ __forceinline__ void store_row_bf16(__bf16 *tile, int q, const float *x) {
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        bf16x8 v;
#pragma unroll
        for (int e = 0; e < 8; ++e) v[e] = to_bf16(x[8 * j + e]);
        *reinterpret_cast<bf16x8 *>(tile + q * kGRow + 8 * j) = v;
    }
}

// guide_encode followed by store_row_bf16 without the 64-float array: the
// row is written eight values at a time (an axis' bands 0-3, then 4-7), so
// only a few of them are live at once. Zeros for a sample without a query.
__device__ __forceinline__ void encode_row_bf16(__bf16 *tile, int q, bool want, const float pos01[3], v3 n,
                                                float lam01, float rough) {
    bf16x8 *row = reinterpret_cast<bf16x8 *>(tile + q * kGRow);
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        float sn = 0.f, cs = 0.f;
        if (want) sincospif(pos01[a], &sn, &cs);
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            bf16x8 v;
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                v[2 * k] = to_bf16(sn);
                v[2 * k + 1] = to_bf16(cs);
                const float s2 = 2.f * sn * cs;
                cs = (cs - sn) * (cs + sn);
                sn = s2;
            }
            row[2 * a + h] = v;
        }
    }
    bf16x8 t0, t1;
#pragma unroll
    for (int e = 0; e < 8; ++e) t0[e] = t1[e] = to_bf16(0.f);
    if (want) {
        t0[0] = to_bf16(n.x);
        t0[1] = to_bf16(n.y);
        t0[2] = to_bf16(n.z);
        t0[3] = to_bf16(lam01);
        t0[4] = to_bf16(rough);
    }
    row[6] = t0;
    row[7] = t1;
}

// One hidden layer for the wave's 64 samples: Hout[q][o] = relu(sum_i W[o][i] Hin[q][i] + b[o]),
// W global bf16 [64][64], Hin/Hout LDS [64][kGRow]. With HoutT non-null the
// transposed copy HoutT[o][q] is written too (for the weight gradient).
__device__ __forceinline__ void layer_fwd(const __bf16 *W, const float *b, const __bf16 *Hin, __bf16 *Hout,
                                          __bf16 *HoutT, bool relu) {
    f32x4 acc[4][4];
    acc_zero(acc);
    mm64(W, kGH, Hin, kGRow, acc);
#pragma unroll
    for (int mt = 0; mt < 4; ++mt)
#pragma unroll
        for (int nt = 0; nt < 4; ++nt) {
            const int q = acc_n(nt);
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int o = acc_m(mt, r);
                float v = acc[mt][nt][r] + b[o];
                if (relu) v = fmaxf(v, 0.f);
                const __bf16 h = to_bf16(v);
                Hout[q * kGRow + o] = h;
                if (HoutT) HoutT[o * kGRow + q] = h;
            }
        }
}

// layer_fwd with the weights at row stride ws (LDS or global) and the
// output written four units at a time (each lane's 4 consecutive units of
// one sample: one 8-byte store). Hout may be Hin: every fragment is loaded
// before the first store, and a wave's LDS operations complete in order.
__device__ __forceinline__ void layer_fwd_rows(const __bf16 *W, int ws, const float *b, const __bf16 *Hin,
                                               __bf16 *Hout) {
    f32x4 acc[4][4];
    acc_zero(acc);
    mm64(W, ws, Hin, kGRow, acc);
#pragma unroll
    for (int mt = 0; mt < 4; ++mt) {
        const int o0 = acc_m(mt, 0);
        const float b0 = b[o0], b1 = b[o0 + 1], b2 = b[o0 + 2], b3 = b[o0 + 3];
#pragma unroll
        for (int nt = 0; nt < 4; ++nt) {
            typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));
            bf16x4 v;
            v[0] = to_bf16(fmaxf(acc[mt][nt][0] + b0, 0.f));
            v[1] = to_bf16(fmaxf(acc[mt][nt][1] + b1, 0.f));
            v[2] = to_bf16(fmaxf(acc[mt][nt][2] + b2, 0.f));
            v[3] = to_bf16(fmaxf(acc[mt][nt][3] + b3, 0.f));
            *reinterpret_cast<bf16x4 *>(Hout + acc_n(nt) * kGRow + o0) = v;
        }
    }
}

// One hidden layer for samples [32 h, 32 h + 32) of a wave's tile H
// ([64][kGRow] bf16), in place: H[q][o] = relu(sum_i W[o][i] H[q][i] + b[o]).
// The half reads and writes only its own 32 rows, so the two halves may run
// one after the other on the same tile; it holds 32 accumulators instead of
// the 64 of a whole layer (for kernels with little register room to spare).
__device__ __forceinline__ void layer_fwd_half(const __bf16 *W, int ws, const float *b, __bf16 *H, int h) {
    f32x4 acc[4][2];
    acc_zero_t(acc);
    mm_tiles<4, 2>(W, ws, H + 32 * h * kGRow, kGRow, acc);
#pragma unroll
    for (int mt = 0; mt < 4; ++mt) {
        const int o0 = acc_m(mt, 0);
        const float b0 = b[o0], b1 = b[o0 + 1], b2 = b[o0 + 2], b3 = b[o0 + 3];
#pragma unroll
        for (int nt = 0; nt < 2; ++nt) {
            typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));
            bf16x4 v;
            v[0] = to_bf16(fmaxf(acc[mt][nt][0] + b0, 0.f));
            v[1] = to_bf16(fmaxf(acc[mt][nt][1] + b1, 0.f));
            v[2] = to_bf16(fmaxf(acc[mt][nt][2] + b2, 0.f));
            v[3] = to_bf16(fmaxf(acc[mt][nt][3] + b3, 0.f));
            *reinterpret_cast<bf16x4 *>(H + (32 * h + acc_n(nt)) * kGRow + o0) = v;
        }
    }
}

// One hidden layer for a wave's 64 samples, in place in the tile H
// ([64][kGRow] bf16): H[q][o] = relu(sum_i W[o][i] H[q][i] + b[o]), with the
// weights read once. All of H's fragments are first loaded into registers
// (32 values per lane), so the tile may be overwritten; the 64 output units
// are then computed in two halves of 32 (32 accumulators live), each reading
// only its own weight rows.
__device__ __forceinline__ void layer_fwd_inreg(const __bf16 *W, int ws, const float *b, __bf16 *H) {
    const int l = threadIdx.x & 63;
    const int r16 = l & 15, kq = (l >> 4) * 8;
    bf16x8 hq[2][4];
#pragma unroll
    for (int ks = 0; ks < 2; ++ks)
#pragma unroll
        for (int nt = 0; nt < 4; ++nt)
            hq[ks][nt] = *reinterpret_cast<const bf16x8 *>(H + (nt * 16 + r16) * kGRow + ks * 32 + kq);
#pragma unroll
    for (int mh = 0; mh < 2; ++mh) {
        f32x4 acc[2][4];
        acc_zero_t(acc);
#pragma unroll
        for (int ks = 0; ks < 2; ++ks) {
            bf16x8 wa[2];
#pragma unroll
            for (int t = 0; t < 2; ++t)
                wa[t] = *reinterpret_cast<const bf16x8 *>(W + ((2 * mh + t) * 16 + r16) * ws + ks * 32 + kq);
#pragma unroll
            for (int t = 0; t < 2; ++t)
#pragma unroll
                for (int nt = 0; nt < 4; ++nt)
                    acc[t][nt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wa[t], hq[ks][nt], acc[t][nt], 0, 0, 0);
        }
#pragma unroll
        for (int t = 0; t < 2; ++t) {
            const int o0 = acc_m(2 * mh + t, 0);
            const float b0 = b[o0], b1 = b[o0 + 1], b2 = b[o0 + 2], b3 = b[o0 + 3];
#pragma unroll
            for (int nt = 0; nt < 4; ++nt) {
                typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));
                bf16x4 v;
                v[0] = to_bf16(fmaxf(acc[t][nt][0] + b0, 0.f));
                v[1] = to_bf16(fmaxf(acc[t][nt][1] + b1, 0.f));
                v[2] = to_bf16(fmaxf(acc[t][nt][2] + b2, 0.f));
                v[3] = to_bf16(fmaxf(acc[t][nt][3] + b3, 0.f));
                *reinterpret_cast<bf16x4 *>(H + acc_n(nt) * kGRow + o0) = v;
            }
        }
    }
}

// Head: raw[q][o] = sum_i W[o][i] H[q][i] + b[o] in fp32 for o < nout, raw
// row stride rs.
__device__ __forceinline__ void head_fwd(const __bf16 *W, const float *b, const __bf16 *H, float *raw, int rs,
                                         int nout = kGH) {
    f32x4 acc[4][4];
    acc_zero(acc);
    mm64(W, kGH, H, kGRow, acc);
#pragma unroll
    for (int mt = 0; mt < 4; ++mt)
#pragma unroll
        for (int nt = 0; nt < 4; ++nt) {
            const int q = acc_n(nt);
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int o = acc_m(mt, r);
                if (o < nout) raw[q * rs + o] = acc[mt][nt][r] + b[o];
            }
        }
}

// d(-c log p)/d raw for one sample: p = sum_i w_i v_i(dir) over the decoded
// lobes. Returns the loss; g receives 40 values.
FL_HD float guide_head_grad(const float *raw, v3 dir, float c, float *g) {
    VmfLobe L[kGLobes];
    guide_decode(raw, L);
    float v[kGLobes], p = 0.f;
    for (int i = 0; i < kGLobes; ++i) {
        v[i] = vmf_pdf(L[i], dir);
        p += L[i].w * v[i];
    }
    p = fmaxf(p, 1e-30f);
    for (int i = 0; i < kGLobes; ++i) {
        const float *r = raw + 5 * i;
        float *gi = g + 5 * i;
        const float resp = L[i].w * v[i] / p;
        // weight logit (softmax)
        gi[4] = -c * (resp - L[i].w);
        // concentration: kappa = exp(clamp(raw)), dlog v/dkappa = 1/kappa - 2/expm1(2 kappa) + (mu.w - 1)
        const float k = L[i].kappa;
        const float md = dot(L[i].mu, dir);
        const bool inside = r[3] > kGKappaLo && r[3] < kGKappaHi;
        const float dk = 1.f / k - 2.f / expm1f(2.f * k) + (md - 1.f);
        gi[3] = inside ? -c * resp * dk * k : 0.f;
        // mean direction mu = m / |m|: dlog v/dm = kappa (w - (mu.w) mu) / |m|
        const float len = sqrtf(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
        if (len > 1e-6f) {
            const float s = -c * resp * k / len;
            gi[0] = s * (dir.x - md * L[i].mu.x);
            gi[1] = s * (dir.y - md * L[i].mu.y);
            gi[2] = s * (dir.z - md * L[i].mu.z);
        } else {
            gi[0] = gi[1] = gi[2] = 0.f;
        }
    }
    return -c * logf(p);
}

} // namespace fl
