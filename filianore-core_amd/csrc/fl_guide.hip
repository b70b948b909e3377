// fl_guide.hip -- training and API of the learned directional proposal
// (vMF mixture MLP, bf16 on the matrix cores, fp32 master weights, Adam).
//
// Loss of one sample (direction w drawn while rendering, weight c):
//     l = -c log p(w),  p = sum_i w_i vMF_i(w)
// with c = radiance the path gathered beyond the vertex = beta f cos L_i /
// p_sample: the sample estimate of the KL divergence from the product
// f cos L_i to the mixture at the vertex, weighted by the vertex's share of
// the pixel (its throughput beta).
//
// Gradients are summed per 64-sample block into a partial buffer and reduced
// in a fixed order, so a training step is deterministic.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <new>
#include <vector>

#include <hipcub/hipcub.hpp>

#include "fl_guide_mlp.h"
#include "fl_internal.h"

using flx::set_error;

namespace {

using namespace fl;

constexpr int kRecEnc = 68;      // API record: encoded input 64, direction 3, weight 1
constexpr int kRecRaw = 16;      // render record (see guide_train_on_records)
constexpr float kClipNorm = 10.f;    // gradient norm clip (mean loss)

__device__ __forceinline__ uint32_t hash_u32(uint32_t x) {
    x ^= x >> 16; x *= 0x7feb352du;
    x ^= x >> 15; x *= 0x846ca68bu;
    x ^= x >> 16;
    return x;
}

// Sources of training samples for k_guide_grad: valid(i), the encoded
// inputs [16 j, 16 j + 16) of sample i, and its direction and weight.
// API records: already encoded.
struct EncSrc {
    const float *rec;
    uint32_t n;
    __device__ bool valid(uint32_t i) const { return i < n; }
    __device__ void quarter(uint32_t i, int j, float *x16) const {
        const float *r = rec + (size_t)i * kRecEnc + 16 * j;
        for (int k = 0; k < 16; ++k) x16[k] = r[k];
    }
    __device__ void target(uint32_t i, v3 &dir, float &c) const {
        const float *r = rec + (size_t)i * kRecEnc;
        dir = v3(r[64], r[65], r[66]);
        c = r[67];
    }
};

// Render records, drawn with probability proportional to their weight
// (k_guide_draw) and encoded here. Each drawn sample carries unit weight: the
// batch gradient has the expectation of the weighted gradient over all
// records (up to the constant total weight), and no batch entry is spent on
// the many records whose path gathered nothing beyond the vertex.
struct RawSrc {
    const float *rec;
    const uint32_t *draw;   // this step's record indices, n_batch of them
    uint32_t n_batch;
    __device__ bool valid(uint32_t i) const { return i < n_batch; }
    __device__ void quarter(uint32_t i, int j, float *x16) const {
        const float *R = rec + (size_t)draw[i] * kRecRaw;
        const float pos[3] = {R[0], R[1], R[2]};
        guide_encode_quarter(j, pos, v3(R[3], R[4], R[5]), R[6], R[7], x16);
    }
    __device__ void target(uint32_t i, v3 &dir, float &c) const {
        const float *R = rec + (size_t)draw[i] * kRecRaw;
        dir = v3(R[8], R[9], R[10]);
        c = 1.f;
    }
};

// The record indices of `steps` Adam steps of `batch` draws each, all at
// once (the searches of a step then no longer sit on its critical path):
// draw i of step k takes the first record whose weight prefix sum exceeds a
// hashed uniform in [0, total). Step k hashes with seed0 ^ (step0 + k), the
// guide's step counter at that step.
__global__ void k_guide_draw(const double *cdf, uint32_t n_recs, double total, uint32_t seed0, uint32_t step0,
                             uint32_t steps, uint32_t batch, uint32_t *out) {
    const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (size_t)steps * batch) return;
    const uint32_t k = (uint32_t)(idx / batch), i = (uint32_t)(idx - (size_t)k * batch);
    const uint32_t seed = seed0 ^ (step0 + k);
    const uint32_t h = hash_u32(hash_u32(seed ^ (k * 0x9e3779b9u)) ^ (i * 0x85ebca6bu));
    const double u = ((double)h + 0.5) * (1.0 / 4294967296.0) * total;
    // first record whose prefix sum exceeds u (zero-weight records never are)
    uint32_t lo = 0, hi = n_recs - 1;
    while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (cdf[mid] > u) hi = mid;
        else lo = mid + 1;
    }
    out[idx] = lo;
}

// Clean weight of render record i (0 unless positive and finite), as double
// for the prefix sum.
struct RecWeight {
    const float *rec;
    __host__ __device__ double operator()(uint32_t i) const {
        const float w = rec[(size_t)i * kRecRaw + 14];
        return (w > 0.f && w < 3.0e38f) ? (double)w : 0.0;
    }
};

// Forward + backward of 64 samples per block of 4 waves, 16 samples per
// wave. The layers' activations of all 64 samples stay in LDS, by sample and
// by unit (the weight gradient's operands are sample-contiguous); each wave
// runs the forward and the data gradient of its own 16 samples (16 x 64
// tiles), and the 64 x 64 weight gradient over the block's samples is split
// by rows, 16 per wave. partial receives this block's gradient sums,
// loss_part its loss sum. Summation orders are fixed: a step is
// deterministic.
constexpr int kGradBlock = 256;
template <class Src>
__global__ void __launch_bounds__(kGradBlock) k_guide_grad(const __bf16 *W, const __bf16 *WT, const float *P, Src src,
                                                           float *partial, float *loss_part) {
    __shared__ __attribute__((aligned(16))) __bf16 sA[kGMats][64 * kGRow];    // X, H1..H4 as [sample][unit]
    __shared__ __attribute__((aligned(16))) __bf16 sAT[kGMats][64 * kGRow];   // the same, [unit][sample]
    __shared__ __attribute__((aligned(16))) __bf16 sD[64 * kGRow], sDT[64 * kGRow];
    __shared__ float sR[64 * 65];                                             // head outputs, fp32 deltas
    __shared__ float sRed[4][64], sLoss[4];
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    const int q0 = 16 * w;                       // the wave's first sample
    const uint32_t base = blockIdx.x * 64u;
    // inputs: lane (sample q0 + l % 16, quarter l / 16 of its 64 inputs)
    {
        const int s = q0 + (l & 15), j = l >> 4;
        float x16[16];
        if (src.valid(base + s)) src.quarter(base + s, j, x16);
        else
            for (int e = 0; e < 16; ++e) x16[e] = 0.f;
        bf16x8 v0, v1;
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            v0[e] = to_bf16(x16[e]);
            v1[e] = to_bf16(x16[8 + e]);
        }
        *reinterpret_cast<bf16x8 *>(sA[0] + s * kGRow + 16 * j) = v0;
        *reinterpret_cast<bf16x8 *>(sA[0] + s * kGRow + 16 * j + 8) = v1;
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            sAT[0][(16 * j + e) * kGRow + s] = v0[e];
            sAT[0][(16 * j + 8 + e) * kGRow + s] = v1[e];
        }
    }
    __syncthreads();
    // forward of the wave's samples: H[q][o] = relu(sum_i W[o][i] H[q][i] + b[o])
    const int qn = q0 + (l & 15);                // this lane's sample in the accumulators
    for (int k = 0; k < 4; ++k) {
        f32x4 acc[4][1];
        acc_zero_t(acc);
        mm_tiles<4, 1>(W + k * 4096, kGH, sA[k] + q0 * kGRow, kGRow, acc);
        const float *b = P + k * kGLayerParams + 4096;
#pragma unroll
        for (int mt = 0; mt < 4; ++mt)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int o = acc_m(mt, r);
                const __bf16 h = to_bf16(fmaxf(acc[mt][0][r] + b[o], 0.f));
                sA[k + 1][qn * kGRow + o] = h;
                sAT[k + 1][o * kGRow + qn] = h;
            }
        __syncthreads();
    }
    {   // head: 40 outputs (three 16-row tiles)
        f32x4 acc[3][1];
        acc_zero_t(acc);
        mm_tiles<3, 1>(W + 4 * 4096, kGH, sA[4] + q0 * kGRow, kGRow, acc);
        const float *b = P + 4 * kGLayerParams + 4096;
#pragma unroll
        for (int mt = 0; mt < 3; ++mt)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int o = acc_m(mt, r);
                if (o < kGOut) sR[qn * 65 + o] = acc[mt][0][r] + b[o];
            }
    }
    __syncthreads();
    // head gradient: lane l < 16 owns sample q0 + l
    float loss = 0.f;
    if (l < 16) {
        const int q = q0 + l;
        float g[kGOut];
        v3 dir(0.f, 0.f, 1.f);
        float c = 0.f;
        const bool ok = src.valid(base + q);
        if (ok) src.target(base + q, dir, c);
        if (ok && c > 0.f) loss = guide_head_grad(sR + q * 65, dir, c, g);
        else
            for (int o = 0; o < kGOut; ++o) g[o] = 0.f;
        for (int o = 0; o < 64; ++o) {
            const float v = o < kGOut ? g[o] : 0.f;
            sR[q * 65 + o] = v;
            const __bf16 h = to_bf16(v);
            sD[q * kGRow + o] = h;
            sDT[o * kGRow + q] = h;
        }
    }
    __syncthreads();
    float *pb = partial + (size_t)blockIdx.x * kGParams;
    for (int k = kGMats - 1; k >= 0; --k) {
        // bias: column l of the fp32 deltas, the wave's 16 samples
        float sb = 0.f;
        for (int s = 0; s < 16; ++s) sb += sR[(q0 + s) * 65 + l];
        sRed[w][l] = sb;
        {   // weight rows [q0, q0 + 16): dW[o][i] = sum_q D[q][o] A[q][i] over the block's samples
            f32x4 acc[1][4];
            acc_zero_t(acc);
            mm_tiles<1, 4>(sDT + q0 * kGRow, kGRow, sAT[k], kGRow, acc);
#pragma unroll
            for (int nt = 0; nt < 4; ++nt)
#pragma unroll
                for (int r = 0; r < 4; ++r) pb[k * kGLayerParams + (q0 + acc_m(0, r)) * 64 + acc_n(nt)] = acc[0][nt][r];
        }
        __syncthreads();   // every wave's weight rows read D^T; the bias parts are in
        if (w == 0) pb[k * kGLayerParams + 4096 + l] = (sRed[0][l] + sRed[1][l]) + (sRed[2][l] + sRed[3][l]);
        if (k == 0) break;
        // data gradient of the wave's samples: dA[q][i] = sum_o WT[i][o] D[q][o],
        // masked by the ReLU of A = relu(Z)
        f32x4 acc[4][1];
        acc_zero_t(acc);
        mm_tiles<4, 1>(WT + k * 4096, kGH, sD + q0 * kGRow, kGRow, acc);
#pragma unroll
        for (int mt = 0; mt < 4; ++mt)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int i = acc_m(mt, r);
                const float a = (float)sA[k][qn * kGRow + i];
                const float d = a > 0.f ? acc[mt][0][r] : 0.f;
                sR[qn * 65 + i] = d;
                const __bf16 h = to_bf16(d);
                sD[qn * kGRow + i] = h;
                sDT[i * kGRow + qn] = h;
            }
        __syncthreads();
    }
    for (int off = 32; off > 0; off >>= 1) loss += __shfl_xor(loss, off);
    if (l == 0) sLoss[w] = loss;
    __syncthreads();
    if (threadIdx.x == 0) loss_part[blockIdx.x] = (sLoss[0] + sLoss[1]) + (sLoss[2] + sLoss[3]);
}

// grad[p] = sum over blocks of partial[b][p] in a fixed order. A block owns
// 64 parameters; its 4 waves each sum a quarter of the partials (lane =
// parameter, coalesced rows), then combine in LDS. sq[blk] receives the sum
// of squares of the block's 64 sums (for the gradient norm).
constexpr int kRedParams = 64;
constexpr int kRedBlocks = (kGParams + kRedParams - 1) / kRedParams;   // 325

__global__ void __launch_bounds__(256) k_grad_reduce(const float *partial, uint32_t nblocks, float *grad, float *sq) {
    __shared__ float red[4][kRedParams];
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    const uint32_t p = blockIdx.x * kRedParams + l;
    float s = 0.f;
    if (p < (uint32_t)kGParams) {
        const uint32_t per = (nblocks + 3) / 4, b0 = w * per, b1 = min(nblocks, b0 + per);
        float s0 = 0.f, s1 = 0.f, s2 = 0.f, s3 = 0.f;
        uint32_t b = b0;
        for (; b + 4 <= b1; b += 4) {
            s0 += partial[(size_t)b * kGParams + p];
            s1 += partial[(size_t)(b + 1) * kGParams + p];
            s2 += partial[(size_t)(b + 2) * kGParams + p];
            s3 += partial[(size_t)(b + 3) * kGParams + p];
        }
        for (; b < b1; ++b) s0 += partial[(size_t)b * kGParams + p];
        s = (s0 + s1) + (s2 + s3);
    }
    red[w][l] = s;
    __syncthreads();
    if (w == 0) {
        const float t = (red[0][l] + red[1][l]) + (red[2][l] + red[3][l]);
        float q = 0.f;
        if (p < (uint32_t)kGParams) {
            grad[p] = t;
            q = t * t;
        }
        for (int off = 32; off > 0; off >>= 1) q += __shfl_xor(q, off);
        if (l == 0) sq[blockIdx.x] = q;
    }
}

// sum of n floats by one 256-thread block, fixed order
__device__ __forceinline__ float block_sum_array(const float *v, uint32_t n, float *red) {
    float s = 0.f;
    for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) s += v[i];
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    if (l == 0) red[w] = s;
    __syncthreads();
    const float t = (red[0] + red[1]) + (red[2] + red[3]);
    __syncthreads();
    return t;
}

// bf16 copies of the master weights: W[k][o][i] and WT[k][i][o]
__device__ __forceinline__ void publish_one(const float *P, __bf16 *W, __bf16 *WT, uint32_t e) {
    const uint32_t k = e / 4096u, oi = e % 4096u, o = oi / 64u, i = oi % 64u;
    const __bf16 h = to_bf16(P[k * kGLayerParams + oi]);
    W[k * 4096u + o * 64u + i] = h;
    WT[k * 4096u + i * 64u + o] = h;
}

__global__ void k_guide_publish(const float *P, __bf16 *W, __bf16 *WT) {
    const uint32_t e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e < (uint32_t)(kGMats * 4096)) publish_one(P, W, WT, e);
}

// One Adam step from the summed gradient: every block recomputes the loss and
// the gradient norm from the partial sums (the same fixed order in every
// block), then updates its 256 parameters and their bf16 copies. The update
// is skipped, with the diverged flag raised, when either is not finite; a
// parameter that is not finite after its update raises the flag too.
__global__ void __launch_bounds__(256) k_adam(float *P, float *M, float *V, const float *grad, const float *sq,
                                              const float *loss_part, uint32_t nblocks, float inv_n, float lr,
                                              float b1, float b2, float eps, uint32_t t, float *stat, __bf16 *W,
                                              __bf16 *WT) {
    __shared__ float red[4];
    const float loss = block_sum_array(loss_part, nblocks, red) * inv_n;
    const float norm = sqrtf(block_sum_array(sq, kRedBlocks, red)) * inv_n;
    if (!isfinite(loss) || !isfinite(norm)) {
        if (blockIdx.x == 0 && threadIdx.x == 0) { stat[0] = loss; stat[1] = norm; stat[2] = 1.f; }
        return;
    }
    const float scale = norm > kClipNorm ? kClipNorm / norm : 1.f;
    const float c1 = 1.f - powf(b1, (float)t), c2 = 1.f - powf(b2, (float)t);
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p < (uint32_t)kGParams) {
        const float g = grad[p] * inv_n * scale;
        const float m = b1 * M[p] + (1.f - b1) * g;
        const float v = b2 * V[p] + (1.f - b2) * g * g;
        M[p] = m;
        V[p] = v;
        const float w = P[p] - lr * (m / c1) / (sqrtf(v / c2) + eps);
        P[p] = w;
        if (!isfinite(w)) stat[2] = 1.f;   // a non-finite parameter (e.g. set from outside) stops training
        const uint32_t k = p / (uint32_t)kGLayerParams, r = p % (uint32_t)kGLayerParams;
        if (r < 4096u) {
            const __bf16 h = to_bf16(w);
            W[k * 4096u + r] = h;
            WT[k * 4096u + (r % 64u) * 64u + r / 64u] = h;
        }
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) { stat[0] = loss; stat[1] = norm; }
}

// Forward of encoded inputs (n x 64) to raw head outputs (n x 40).
__global__ void __launch_bounds__(64) k_guide_fwd(const __bf16 *W, const float *P, const float *in, float *out,
                                                  uint32_t n) {
    __shared__ __bf16 sH[2][64 * kGRow];
    __shared__ float sR[64 * 65];
    const int q = threadIdx.x;
    const uint32_t s = blockIdx.x * 64u + q;
    for (int i = 0; i < 64; ++i) sH[0][q * kGRow + i] = to_bf16(s < n ? in[(size_t)s * 64 + i] : 0.f);
    __syncthreads();
    for (int k = 0; k < 4; ++k) {
        layer_fwd(W + k * 4096, P + k * kGLayerParams + 4096, sH[k & 1], sH[(k + 1) & 1], nullptr, true);
        __syncthreads();
    }
    head_fwd(W + 4 * 4096, P + 4 * kGLayerParams + 4096, sH[0], sR, 65);
    __syncthreads();
    if (s < n)
        for (int o = 0; o < kGOut; ++o) out[(size_t)s * kGOut + o] = sR[q * 65 + o];
}

__device__ __forceinline__ float block_sum(float v, float *red) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    if (l == 0) red[w] = v;
    __syncthreads();
    float t = 0.f;
    if (threadIdx.x == 0) {
        for (int i = 0; i < (int)(blockDim.x >> 6); ++i) t += red[i];
        red[32] = t;
    }
    __syncthreads();
    t = red[32];
    __syncthreads();
    return t;
}

// ---------------------------------------------------------------- host ----
struct Pcg {
    uint64_t s, inc;
    explicit Pcg(uint64_t seed) : s(0), inc(0xda3e39cb94b95bdbULL) { next(); s += seed; next(); }
    uint32_t next() {
        uint64_t o = s;
        s = o * 6364136223846793005ULL + inc;
        uint32_t x = (uint32_t)(((o >> 18u) ^ o) >> 27u), r = (uint32_t)(o >> 59u);
        return (x >> r) | (x << ((-r) & 31));
    }
    float uniform() { return (float)(next() >> 8) * (1.f / 16777216.f); }
};

// He-uniform hidden layers; the head starts small with its lobe means on the
// 8 cube diagonals, concentration 1 and equal weights.
void init_params(std::vector<float> &P, uint64_t seed) {
    P.assign(kGParams, 0.f);
    Pcg rng(seed);
    for (int k = 0; k < kGMats; ++k) {
        const int fan_in = k == 0 ? kGIn : kGH;
        const float a = k < 4 ? std::sqrt(6.f / (float)fan_in) : 0.1f * std::sqrt(6.f / (float)fan_in);
        const int rows = k < 4 ? kGH : kGOut;
        for (int o = 0; o < rows; ++o)
            for (int i = 0; i < fan_in; ++i) P[k * kGLayerParams + o * 64 + i] = a * (2.f * rng.uniform() - 1.f);
    }
    float *b4 = P.data() + 4 * kGLayerParams + 4096;
    for (int l = 0; l < kGLobes; ++l) {
        b4[5 * l + 0] = (l & 1) ? 1.f : -1.f;
        b4[5 * l + 1] = (l & 2) ? 1.f : -1.f;
        b4[5 * l + 2] = (l & 4) ? 1.f : -1.f;
    }
}

int ensure(fl_ctx_t ctx, float *&p, size_t &cap, size_t bytes) {
    if (cap >= bytes) return FL_OK;
    if (p) (void)hipFree(p);
    p = nullptr;
    cap = 0;
    if (hipMalloc(&p, bytes) != hipSuccess) return set_error(ctx, FL_E_OOM, "guide staging buffer");
    cap = bytes;
    return FL_OK;
}

template <class T>
int ensure_bytes(fl_ctx_t ctx, T *&p, size_t &cap, size_t bytes) {
    if (cap >= bytes) return FL_OK;
    if (p) (void)hipFree(p);
    p = nullptr;
    cap = 0;
    if (hipMalloc(reinterpret_cast<void **>(&p), bytes) != hipSuccess)
        return set_error(ctx, FL_E_OOM, "guide record buffer");
    cap = bytes;
    return FL_OK;
}

int ensure_partial(fl_guide_s *g, uint32_t blocks) {
    if (g->partial_blocks >= blocks) return FL_OK;
    if (g->d_partial) (void)hipFree(g->d_partial);
    if (g->d_loss) (void)hipFree(g->d_loss);
    g->d_partial = nullptr;
    g->d_loss = nullptr;
    g->partial_blocks = 0;
    if (hipMalloc(&g->d_partial, (size_t)blocks * kGParams * 4) != hipSuccess ||
        hipMalloc(&g->d_loss, (size_t)blocks * 4) != hipSuccess)
        return set_error(g->ctx, FL_E_OOM, "guide gradient buffers");
    g->partial_blocks = blocks;
    return FL_OK;
}

// summed gradient of n samples into g->d_grad, loss sum partials in d_loss
template <class Src>
int run_grad(fl_guide_s *g, const Src &src, uint32_t n, hipStream_t st) {
    const uint32_t blocks = (n + 63) / 64;
    int rc = ensure_partial(g, blocks);
    if (rc != FL_OK) return rc;
    hipLaunchKernelGGL(k_guide_grad<Src>, dim3(blocks), dim3(kGradBlock), 0, st, g->d_w, g->d_wt, g->d_params, src,
                       g->d_partial, g->d_loss);
    hipLaunchKernelGGL(k_grad_reduce, dim3(kRedBlocks), dim3(256), 0, st, g->d_partial, blocks, g->d_grad, g->d_sq);
    return FL_OK;
}

void adam(fl_guide_s *g, uint32_t blocks, uint32_t n, hipStream_t st) {
    g->steps += 1;
    hipLaunchKernelGGL(k_adam, dim3((kGParams + 255) / 256), dim3(256), 0, st, g->d_params, g->d_m, g->d_v, g->d_grad,
                       g->d_sq, g->d_loss, blocks, 1.f / (float)std::max(1u, n), g->cfg.lr, g->cfg.beta1,
                       g->cfg.beta2, g->cfg.eps, g->steps, g->d_stat, g->d_w, g->d_wt);
}

int read_stats(fl_guide_s *g) {
    float h[3];
    if (hipMemcpyAsync(h, g->d_stat, sizeof(h), hipMemcpyDeviceToHost, g->ctx->stream) != hipSuccess ||
        hipStreamSynchronize(g->ctx->stream) != hipSuccess)
        return set_error(g->ctx, FL_E_HIP, "guide stats");
    g->last_loss = h[0];
    if (h[2] != 0.f) g->diverged = 1;
    return FL_OK;
}

} // namespace

namespace flx {

int guide_train_on_records(fl_guide_s *g, const float *d_recs, uint32_t n_recs, uint32_t steps, uint32_t batch,
                           uint64_t seed, hipStream_t st) {
    if (!g || n_recs == 0 || steps == 0 || g->diverged) return FL_OK;
    // prefix sum of the record weights, for weight-proportional draws
    int rc = ensure_bytes(g->ctx, g->d_cdf, g->cdf_bytes, (size_t)n_recs * sizeof(double));
    if (rc != FL_OK) return rc;
    hipcub::TransformInputIterator<double, RecWeight, hipcub::CountingInputIterator<uint32_t>> wit(
        hipcub::CountingInputIterator<uint32_t>(0u), RecWeight{d_recs});
    size_t tmp_bytes = 0;
    if (hipcub::DeviceScan::InclusiveSum(nullptr, tmp_bytes, wit, g->d_cdf, n_recs, st) != hipSuccess)
        return set_error(g->ctx, FL_E_HIP, "guide record scan");
    rc = ensure_bytes(g->ctx, g->d_scan_tmp, g->scan_tmp_bytes, std::max<size_t>(tmp_bytes, 256));
    if (rc != FL_OK) return rc;
    if (hipcub::DeviceScan::InclusiveSum(g->d_scan_tmp, tmp_bytes, wit, g->d_cdf, n_recs, st) != hipSuccess)
        return set_error(g->ctx, FL_E_HIP, "guide record scan");
    double total = 0.0;
    if (hipMemcpyAsync(&total, g->d_cdf + (n_recs - 1), sizeof(double), hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess)
        return set_error(g->ctx, FL_E_HIP, "guide record weights");
    if (!(total > 0.0) || !std::isfinite(total)) return FL_OK;   // nothing to learn from yet
    const uint32_t blocks = (batch + 63) / 64;
    // every step's draws at once
    const size_t n_draw = (size_t)steps * batch;
    rc = ensure_bytes(g->ctx, g->d_draw, g->draw_bytes, n_draw * sizeof(uint32_t));
    if (rc != FL_OK) return rc;
    hipLaunchKernelGGL(k_guide_draw, dim3((unsigned)((n_draw + 255) / 256)), dim3(256), 0, st, g->d_cdf, n_recs, total,
                       (uint32_t)(seed * 0x9e3779b97f4a7c15ULL >> 32), g->steps, steps, batch, g->d_draw);
    for (uint32_t k = 0; k < steps; ++k) {
        RawSrc src{d_recs, g->d_draw + (size_t)k * batch, batch};
        rc = run_grad(g, src, batch, st);
        if (rc != FL_OK) return rc;
        adam(g, blocks, batch, st);
    }
    if (hipGetLastError() != hipSuccess) return set_error(g->ctx, FL_E_HIP, "guide training launch");
    return read_stats(g);   // synchronises: the caller's timing covers the steps
}

} // namespace flx

extern "C" {

int fl_guide_create(fl_ctx_t ctx, const fl_guide_config *cfg, fl_guide_t *out) {
    if (!ctx || !out) return set_error(ctx, FL_E_INVALID, "fl_guide_create: null argument");
    *out = nullptr;
    fl_guide_config c{};
    if (cfg) c = *cfg;
    if ((c.hidden && c.hidden != (uint32_t)kGH) || (c.layers && c.layers != 4u) ||
        (c.lobes && c.lobes != (uint32_t)kGLobes) || (c.bands && c.bands != (uint32_t)kGBands))
        return set_error(ctx, FL_E_UNSUPPORTED, "fl_guide_create: the network is 4 x 64 hidden, 8 lobes, 8 bands");
    c.hidden = kGH;
    c.layers = 4;
    c.lobes = kGLobes;
    c.bands = kGBands;
    if (!(c.lr > 0.f)) c.lr = 5e-3f;
    if (!(c.beta1 > 0.f)) c.beta1 = 0.9f;
    if (!(c.beta2 > 0.f)) c.beta2 = 0.99f;
    if (!(c.eps > 0.f)) c.eps = 1e-8f;
    if (!c.batch) c.batch = 16384;
    if (!c.steps_per_pass) c.steps_per_pass = 128;
    if (!c.train_rounds) c.train_rounds = 3;
    if (c.train_rounds > 16) return set_error(ctx, FL_E_INVALID, "fl_guide_create: train_rounds above 16");
    fl_guide_s *g = new (std::nothrow) fl_guide_s();
    if (!g) return set_error(ctx, FL_E_OOM, "fl_guide_create");
    g->ctx = ctx;
    g->cfg = c;
    (void)hipSetDevice(ctx->device);
    bool ok = hipMalloc(&g->d_params, kGParams * 4) == hipSuccess && hipMalloc(&g->d_m, kGParams * 4) == hipSuccess &&
              hipMalloc(&g->d_v, kGParams * 4) == hipSuccess && hipMalloc(&g->d_grad, kGParams * 4) == hipSuccess &&
              hipMalloc(&g->d_w, kGMats * 4096 * 2) == hipSuccess &&
              hipMalloc(&g->d_wt, kGMats * 4096 * 2) == hipSuccess && hipMalloc(&g->d_stat, 16) == hipSuccess &&
              hipMalloc(&g->d_sq, kRedBlocks * 4) == hipSuccess;
    if (!ok) {
        fl_guide_destroy(g);
        return set_error(ctx, FL_E_OOM, "fl_guide_create: device buffers");
    }
    std::vector<float> P;
    init_params(P, c.seed ? c.seed : 1u);
    std::vector<float> z(kGParams, 0.f);
    ok = hipMemcpy(g->d_params, P.data(), kGParams * 4, hipMemcpyHostToDevice) == hipSuccess &&
         hipMemcpy(g->d_m, z.data(), kGParams * 4, hipMemcpyHostToDevice) == hipSuccess &&
         hipMemcpy(g->d_v, z.data(), kGParams * 4, hipMemcpyHostToDevice) == hipSuccess &&
         hipMemcpy(g->d_stat, z.data(), 16, hipMemcpyHostToDevice) == hipSuccess;
    if (ok) {
        hipLaunchKernelGGL(k_guide_publish, dim3((kGMats * 4096 + 255) / 256), dim3(256), 0, ctx->stream, g->d_params,
                           g->d_w, g->d_wt);
        ok = hipStreamSynchronize(ctx->stream) == hipSuccess && hipGetLastError() == hipSuccess;
    }
    if (!ok) {
        fl_guide_destroy(g);
        return set_error(ctx, FL_E_HIP, "fl_guide_create: upload");
    }
    *out = g;
    return FL_OK;
}

void fl_guide_destroy(fl_guide_t g) {
    if (!g) return;
    if (g->ctx) (void)hipStreamSynchronize(g->ctx->stream);
    for (void *p : {(void *)g->d_params, (void *)g->d_m, (void *)g->d_v, (void *)g->d_grad, (void *)g->d_sq, (void *)g->d_w,
                    (void *)g->d_wt, (void *)g->d_partial, (void *)g->d_loss, (void *)g->d_stat,
                    (void *)g->d_scratch_in, (void *)g->d_scratch_out, (void *)g->d_cdf, g->d_scan_tmp,
                    (void *)g->d_draw})
        if (p) (void)hipFree(p);
    delete g;
}

int fl_guide_forward(fl_guide_t g, const float *inputs, float *outputs, uint32_t n) {
    if (!g || (!inputs && n) || (!outputs && n)) return set_error(g ? g->ctx : nullptr, FL_E_INVALID, "fl_guide_forward");
    if (n == 0) return FL_OK;
    fl_ctx_t ctx = g->ctx;
    int rc = ensure(ctx, g->d_scratch_in, g->scratch_in_bytes, (size_t)n * 64 * 4);
    if (rc == FL_OK) rc = ensure(ctx, g->d_scratch_out, g->scratch_out_bytes, (size_t)n * kGOut * 4);
    if (rc != FL_OK) return rc;
    if (hipMemcpy(g->d_scratch_in, inputs, (size_t)n * 64 * 4, hipMemcpyHostToDevice) != hipSuccess)
        return set_error(ctx, FL_E_HIP, "fl_guide_forward: upload");
    hipLaunchKernelGGL(k_guide_fwd, dim3((n + 63) / 64), dim3(64), 0, ctx->stream, g->d_w, g->d_params,
                       g->d_scratch_in, g->d_scratch_out, n);
    if (hipMemcpyAsync(outputs, g->d_scratch_out, (size_t)n * kGOut * 4, hipMemcpyDeviceToHost, ctx->stream) !=
            hipSuccess ||
        hipStreamSynchronize(ctx->stream) != hipSuccess || hipGetLastError() != hipSuccess)
        return set_error(ctx, FL_E_HIP, "fl_guide_forward");
    return FL_OK;
}

int fl_guide_gradient(fl_guide_t g, const float *records, uint32_t n, float *grad, uint32_t n_params, float *loss) {
    if (!g || !records || !grad || n == 0) return set_error(g ? g->ctx : nullptr, FL_E_INVALID, "fl_guide_gradient");
    if (n_params < (uint32_t)kGParams) return set_error(g->ctx, FL_E_INVALID, "fl_guide_gradient: grad too small");
    fl_ctx_t ctx = g->ctx;
    int rc = ensure(ctx, g->d_scratch_in, g->scratch_in_bytes, (size_t)n * kRecEnc * 4);
    if (rc != FL_OK) return rc;
    if (hipMemcpy(g->d_scratch_in, records, (size_t)n * kRecEnc * 4, hipMemcpyHostToDevice) != hipSuccess)
        return set_error(ctx, FL_E_HIP, "fl_guide_gradient: upload");
    rc = run_grad(g, EncSrc{g->d_scratch_in, n}, n, ctx->stream);
    if (rc != FL_OK) return rc;
    const uint32_t blocks = (n + 63) / 64;
    std::vector<float> lp(blocks);
    if (hipMemcpyAsync(grad, g->d_grad, kGParams * 4, hipMemcpyDeviceToHost, ctx->stream) != hipSuccess ||
        hipMemcpyAsync(lp.data(), g->d_loss, blocks * 4, hipMemcpyDeviceToHost, ctx->stream) != hipSuccess ||
        hipStreamSynchronize(ctx->stream) != hipSuccess || hipGetLastError() != hipSuccess)
        return set_error(ctx, FL_E_HIP, "fl_guide_gradient");
    if (loss) {
        double s = 0.0;
        for (float v : lp) s += v;
        *loss = (float)s;
    }
    return FL_OK;
}

int fl_guide_train_step(fl_guide_t g, const float *records, uint32_t n, int records_on_device, float *loss) {
    if (!g || !records || n == 0) return set_error(g ? g->ctx : nullptr, FL_E_INVALID, "fl_guide_train_step");
    fl_ctx_t ctx = g->ctx;
    const float *d_rec = records;
    if (!records_on_device) {
        int rc = ensure(ctx, g->d_scratch_in, g->scratch_in_bytes, (size_t)n * kRecEnc * 4);
        if (rc != FL_OK) return rc;
        if (hipMemcpy(g->d_scratch_in, records, (size_t)n * kRecEnc * 4, hipMemcpyHostToDevice) != hipSuccess)
            return set_error(ctx, FL_E_HIP, "fl_guide_train_step: upload");
        d_rec = g->d_scratch_in;
    }
    if (g->diverged) return set_error(ctx, FL_E_NUMERIC, "guide diverged; training stopped");
    int rc = run_grad(g, EncSrc{d_rec, n}, n, ctx->stream);
    if (rc != FL_OK) return rc;
    adam(g, (n + 63) / 64, n, ctx->stream);
    if (hipStreamSynchronize(ctx->stream) != hipSuccess || hipGetLastError() != hipSuccess)
        return set_error(ctx, FL_E_HIP, "fl_guide_train_step");
    rc = read_stats(g);
    if (rc != FL_OK) return rc;
    if (loss) *loss = g->last_loss;
    if (g->diverged) return set_error(ctx, FL_E_NUMERIC, "guide diverged (non-finite loss or gradient)");
    return FL_OK;
}

int fl_guide_get_params(fl_guide_t g, float *params, uint32_t n_params, uint32_t *n_out) {
    if (!g) return FL_E_INVALID;
    if (n_out) *n_out = 3u * kGParams;
    if (!params) return FL_OK;
    if (n_params < 3u * kGParams) return set_error(g->ctx, FL_E_INVALID, "fl_guide_get_params: buffer too small");
    if (hipMemcpy(params, g->d_params, kGParams * 4, hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(params + kGParams, g->d_m, kGParams * 4, hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(params + 2 * kGParams, g->d_v, kGParams * 4, hipMemcpyDeviceToHost) != hipSuccess)
        return set_error(g->ctx, FL_E_HIP, "fl_guide_get_params");
    return FL_OK;
}

int fl_guide_set_params(fl_guide_t g, const float *params, uint32_t n_params) {
    if (!g || !params || (n_params != (uint32_t)kGParams && n_params != 3u * kGParams))
        return set_error(g ? g->ctx : nullptr, FL_E_INVALID, "fl_guide_set_params: need 20800 or 62400 floats");
    fl_ctx_t ctx = g->ctx;
    if (hipMemcpy(g->d_params, params, kGParams * 4, hipMemcpyHostToDevice) != hipSuccess)
        return set_error(ctx, FL_E_HIP, "fl_guide_set_params");
    if (n_params == 3u * kGParams &&
        (hipMemcpy(g->d_m, params + kGParams, kGParams * 4, hipMemcpyHostToDevice) != hipSuccess ||
         hipMemcpy(g->d_v, params + 2 * kGParams, kGParams * 4, hipMemcpyHostToDevice) != hipSuccess))
        return set_error(ctx, FL_E_HIP, "fl_guide_set_params");
    const float zero[4] = {0.f, 0.f, 0.f, 0.f};
    if (hipMemcpy(g->d_stat, zero, sizeof(zero), hipMemcpyHostToDevice) != hipSuccess)
        return set_error(ctx, FL_E_HIP, "fl_guide_set_params");
    hipLaunchKernelGGL(k_guide_publish, dim3((kGMats * 4096 + 255) / 256), dim3(256), 0, ctx->stream, g->d_params,
                       g->d_w, g->d_wt);
    if (hipStreamSynchronize(ctx->stream) != hipSuccess || hipGetLastError() != hipSuccess)
        return set_error(ctx, FL_E_HIP, "fl_guide_set_params: publish");
    g->diverged = 0;
    return FL_OK;
}

int fl_guide_stats(fl_guide_t g, uint32_t *steps, float *last_loss, int *diverged) {
    if (!g) return FL_E_INVALID;
    if (steps) *steps = g->steps;
    if (last_loss) *last_loss = g->last_loss;
    if (diverged) *diverged = g->diverged;
    return FL_OK;
}

} // extern "C"
