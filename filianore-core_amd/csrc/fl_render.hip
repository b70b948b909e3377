// fl_render.hip -- wavefront spectral path tracer (PathIntegrator::Li for
// every camera sample, integrators/pathintegrator.cpp:12-87, with
// uniform_sample_all_lights/estimate_direct, integrators/integrator.cpp:10-110).
//
// Kernels per pass:   generate -> { trace -> shade } x bounces -> trace -> resolve
//   generate : camera rays (PhysicalCamera::awaken_ray, cameras/physical.cpp:12-22,
//              pinhole), hero wavelengths, path state init
//   trace    : one persistent launch per bounce over the previous bounce's
//              shadow rays (any hit, emitters transparent: VisibilityEvaluator::
//              unoccluded, illuminants/illuminant.cpp:40-50) followed by this
//              bounce's extension rays (closest hit); lanes refill as rays end
//   shade    : emission (+MIS), NEE light samples -> shadow queue, BSDF
//              sample -> next ray queue, depth limits, Russian roulette
//   resolve  : hero lanes -> CIE XYZ, per-pixel sums in a fixed order
// Queues are SoA arrays appended with one atomic per wave (ballot + popcount).
// Queue counters live in device memory and are double-buffered by bounce
// parity, so every kernel reads its work size from the device (the bounce loop
// needs no host round trip and can be captured in a graph).
// The MIS BSDF ray of estimate_direct is the next extension ray (same
// expectation, one traversal fewer per light per vertex).
//
// The file is compiled twice. The main unit holds the traversal, generation
// and resolve kernels and the host driver, with correctly rounded fp32
// division and square root (the hits must equal the CPU's bit for bit). The
// shading unit (FL_SHADE_TU) holds the shading kernels and the BSDF query
// kernels, with the hardware's division and square root (within 2 ulp): a
// third of the shading kernel's instructions were the correctly rounded
// division sequences, and shading the spheres workload took 9% longer with
// them. Shading results are compared with the oracle statistically and with
// FP64 BSDFs to 1e-5, far above that rounding.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>

#include "fl_guide_mlp.h"
#include "fl_wavefront.h"

namespace {

using namespace fl;

constexpr int kBlock = 256;
// device bytes per path slot: state 48, pending NEE 32, two extension queues
// 64, hit 8 (+ 8 of barycentrics in textured renders), shadow queue (two
// entries) 56, and their NEE terms for occluded rays (scenes with shadow
// colors) 32
constexpr size_t kBytesPerPath = 248;

// device counter slots
constexpr int C_EXT = 0;    // [0],[1]: extension queue sizes by bounce parity
constexpr int C_SH = 2;     // [2],[3]: shadow queue sizes by the parity of the bounce that wrote them
constexpr int C_FETCH = 4;  // [4],[5]: trace refill cursors by bounce parity

// Reserves `wave_count` queue slots for each wave of the block with one
// atomic per block and returns the calling wave's first slot. Every thread of
// the block must call it; `s` (kBlock/64 + 1 words of LDS) may be reused by
// the next call only after another barrier of the block.
template <int BS = kBlock>
__device__ __forceinline__ uint32_t block_reserve(uint32_t wave_count, uint32_t *counter, uint32_t *s) {
    constexpr uint32_t kWaves = BS / 64;
    const uint32_t w = threadIdx.x >> 6;
    if (__lane_id() == 0) s[w] = wave_count;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t tot = 0;
        for (uint32_t k = 0; k < kWaves; ++k) {
            const uint32_t c = s[k];
            s[k] = tot;
            tot += c;
        }
        s[kWaves] = tot ? atomicAdd(counter, tot) : 0u;
    }
    __syncthreads();
    return s[kWaves] + s[w];
}

__device__ __forceinline__ v3 ld3(const float *p) { return v3(p[0], p[1], p[2]); }

__global__ void k_zero_u32(uint32_t *p, uint32_t n) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0u;
}
__global__ void k_zero_f32(float *p, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0.f;
}

// ------------------------------------------------------------- generate ----
// Path ids run item after item; inside an item over sample layers of npix
// padded pixels in 8x8 tile order. Queue slots skip the padding. The slot of
// an in-image pixel is its item's slot base plus its rank among the item's
// in-image pixels in the same order, in closed form (no atomics): full tile
// rows before it, full tiles before it in its tile row, then rows and columns
// inside its (possibly clipped) tile.
__global__ void __launch_bounds__(kBlock) k_generate(RenderArgs a) {
    const uint32_t stride = gridDim.x * kBlock;
    if (blockIdx.x == 0 && threadIdx.x == 0) a.counts[C_EXT + 0] = a.n_cam_rays;
    for (uint32_t sid = blockIdx.x * kBlock + threadIdx.x; sid < a.n_paths; sid += stride) {
        const uint32_t item = a.n_items > 1 ? item_search(a, sid, [](const ItemDev &it) { return it.path_base; }) : 0u;
        const ItemDev &it = a.items[item];
        const uint32_t W = it.cam.W, H = it.cam.H;
        uint32_t x = 0, y = 0, pixel = 0, sample = 0, slot = 0;
        const uint32_t local = sid - it.path_base;
        uint32_t s = local / it.npix, pl = local - s * it.npix;
        uint32_t tile = pl >> 6, within = pl & 63u;
        uint32_t tx = tile % it.tiles_x, ty = tile / it.tiles_x, wx = within & 7u, wy = within >> 3;
        x = tx * 8u + wx;
        y = ty * 8u + wy;
        const bool active = x < W && y < H;
        if (active) {
            pixel = y * W + x;
            sample = it.sample_base + s;
            const uint32_t rows_in = min(8u, H - ty * 8u), cols_in = min(8u, W - tx * 8u);
            slot = it.slot_base + s * (W * H) + 8u * W * ty + 8u * tx * rows_in + wy * cols_in + wx;
        }
        v3 o, d;
        if (active) {
            uint32_t rb = rng_base(it.seed, pixel, sample, 0xffu);
            float jx = rng_u(rb, 0), jy = rng_u(rb, 1), ul = rng_u(rb, 2);
            float lam0 = a.grid.lo + ul * (a.grid.hi - a.grid.lo);
            if (lam0 >= a.grid.hi) lam0 = a.grid.lo;
            // pixel column px -> s = 1 - (px+jx)/W, row py (top=0) -> t = 1 - (py+jy)/H
            float su = 1.f - ((float)x + jx) / (float)W;
            float tv = 1.f - ((float)y + jy) / (float)H;
            v3 llc = ld3(it.cam.llc), hor = ld3(it.cam.hor), ver = ld3(it.cam.ver);
            o = ld3(it.cam.o);
            v3 dir = llc + hor * su + ver * tv - o - v3(0.f, 0.f, 0.f);
            d = normalize(dir);
            // throughput, radiance and the NEE slots start at 1, 0, 0: the
            // path's first shade takes them from PF_FRESH instead of memory
            // (64 B per path neither written here nor read there)
            a.st_misc[sid] = make_float4(lam0, 1.f, 0.f, __uint_as_float(PF_FRESH));
        } else {
            a.st_L[sid] = make_float4(0.f, 0.f, 0.f, 0.f);
            a.st_misc[sid] = make_float4(0.f, 1.f, 0.f, __uint_as_float(PF_DONE));
            a.pend[sid] = make_float4(0.f, 0.f, 0.f, 0.f);
            a.pend[a.cap + sid] = make_float4(0.f, 0.f, 0.f, 0.f);
        }
        if (a.n_items > 1) a.st_item[sid] = item;
        if (a.st_rec) a.st_rec[sid] = 0u;
        if (active) {
            a.rq_o[0][slot] = make_float4(o.x, o.y, o.z, 0.f);       // tMin = time = 0 (physical.cpp:21)
            a.rq_d[0][slot] = Packed3{d.x, d.y, d.z};
            a.rq_sid[0][slot] = sid;
        }
    }
}

// ---------------------------------------------------------------- trace ----
// Index space: [0, ns) shadow rays written by the previous bounce's shade,
// [ns, ns + ne) this bounce's extension rays.
// In batches each ray traverses the BVH of its path's item.
//
// Shadow rays follow VisibilityEvaluator::unoccluded (illuminant.cpp:40-50):
// the closest hit on the segment decides, and the segment counts as
// unoccluded when that hit is an emitter. An area light's own segment ends on
// the light, so any hit before it that is not an emitter decides the same
// way: those rays run any-hit with emitters skipped. A delta light's segment
// can pass through an emitter and go on; its shadow rays (target slot tagged
// kVisClosest) run closest-hit and test the hit's light flag.
constexpr uint32_t kVisClosest = 0x80000000u;
// the any-hit word of an extension ray (closest hit), read like a shadow
// target word so that neither kind of ray tests a loaded value before it starts
__device__ const uint32_t kClosestWord = kVisClosest;

// MULTI: the pass holds several items (a batch), each ray traverses its
// path's item's BVH; otherwise every ray traverses sc0, whose pointers stay
// uniform (scalar registers) in the traversal.
template <bool MULTI>
struct TraceSrc {
    static constexpr bool kMulti = MULTI;   // rays of one wave may traverse different BVHs
    const Packed3 *so, *sd;
    const uint32_t *stgt;
    float4 *pend;
    const float4 *eo;
    const Packed3 *ed;
    const uint32_t *esid;
    float2 *hits;            // per extension ray: t and triangle record
    uint32_t ns, cap;
    const ItemDev *items;
    const uint32_t *st_item;
    uint32_t n_items;
    TraceScene sc0;
    const float4 *socc;      // NEE term of an occluded shadow ray (shadowColor), or null: 0
    float2 *huv;             // per extension ray: barycentrics (textured renders; else null)
    __device__ TraceScene base_scene() const { return sc0; }
    __device__ TraceScene scene_of_ray(uint32_t i) const {
        if (!MULTI || n_items <= 1) return sc0;
        uint32_t sid;
        if (i < ns) {
            const uint32_t t = stgt[i] & ~kVisClosest;
            sid = t >= cap ? t - cap : t;   // target slot j * cap + path id
        } else {
            sid = esid[i - ns];
        }
        return items[st_item[sid]].sc;
    }
    __device__ void load(uint32_t i, v3 &ro, v3 &rd, float &tmin, float &tmax, uint32_t &anyw, TraceScene &sc) const {
        float4 a;
        Packed3 b;
        // a shadow ray is any-hit unless its target word carries kVisClosest
        anyw = *(i < ns ? stgt + i : &kClosestWord);
        if (i < ns) {
            const Packed3 o3 = so[i];
            a = make_float4(o3.x, o3.y, o3.z, kShadowEps);   // every shadow ray starts kShadowEps out
            b = sd[i];
        } else {
            a = eo[i - ns];
            b = ed[i - ns];
        }
        sc = scene_of_ray(i);
        ro = v3(a.x, a.y, a.z);
        rd = v3(b.x, b.y, b.z);
        tmin = a.w;
        // the queues' segment ends are fixed: a shadow ray stops SHADOW_EPSILON
        // short of its target (d spans the segment), an extension ray is
        // unbounded; set here rather than loaded, so that a prefetched ray
        // holds one register less
        tmax = i < ns ? 1.f - kShadowEps : kInf;
    }
    __device__ void store(uint32_t i, const TravState &s) const {
        if (i < ns) {
            const uint32_t t = stgt[i];
            bool occluded = s.best_id != kNoHit;
            if (occluded && (t & kVisClosest)) {
                // closest hit on the segment: an emitter there leaves it unoccluded
                const uint4 *tp = reinterpret_cast<const uint4 *>(scene_of_ray(i).tris + s.best_rec);
                occluded = tp[2].w == 0u;
            }
            if (occluded)   // Li = shadowColor (integrator.cpp:47-50)
                pend[t & ~kVisClosest] = socc ? socc[i] : make_float4(0.f, 0.f, 0.f, 0.f);
        } else {
            hits[i - ns] = make_float2(s.best_t, __uint_as_float(s.best_id != kNoHit ? s.best_rec : kNoHit));
            if (huv) huv[i - ns] = make_float2(s.bu, s.bv);
        }
    }
};

// STACK: LDS short-stack entries per lane (deeper entries spill to global
// memory); WAVES: minimum waves per SIMD asked of the register allocator.
template <bool COUNT, int STACK, int WAVES, bool PREFETCH, bool MULTI>
__global__ void __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(WAVES))) k_trace(RenderArgs a, int b) {
    __shared__ unsigned long long stk[STACK * kBlock];
    const int cur = b & 1;
    const uint32_t ns = a.counts[C_SH + (cur ^ 1)];
    const uint32_t ne = a.counts[C_EXT + cur];
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        a.counts[C_EXT + (cur ^ 1)] = 0u;        // shade(b) appends the next extension queue here
        a.counts[C_SH + cur] = 0u;               // and its shadow rays here
        a.counts[C_FETCH + (cur ^ 1)] = 0u;      // refill cursor of trace(b+1)
        atomicAdd(&a.stats[0], (unsigned long long)ne);
        atomicAdd(&a.stats[1], (unsigned long long)ns);
    }
    LaneStack ls;
    ls.lds = (lds_u64 *)(stk + threadIdx.x);
    ls.lds_stride = kBlock;
    ls.cap = STACK;
    ls.spill = a.spill + blockIdx.x * kBlock + threadIdx.x;
    ls.spill_stride = a.spill_stride;
    TravCounters c;
    TraceSrc<MULTI> src{a.sq_o,    a.sq_d, a.sq_tgt, a.pend,    a.rq_o[cur], a.rq_d[cur], a.rq_sid[cur],
                        a.hits,    ns,     a.cap,    a.items,   a.st_item,   a.n_items,   a.sc, a.sq_occ, a.hit_uv};
    // refill threshold by default: a yield once 24 lanes are done when the
    // next rays are loaded ahead (spheres, 64 spp: 48.1 against 48.8 ms at
    // 16 and 52.7 at 8), a refill once 16 lanes are idle otherwise
    const int refill = a.refill_lanes ? a.refill_lanes : (PREFETCH ? 24 : 16);
    trace_persistent<COUNT, true, PREFETCH>(src, ns + ne, &a.counts[C_FETCH + cur], ls, a.trav_budget, refill, c);
    if (COUNT) {
        atomicAdd(&a.stats[3], (unsigned long long)c.nodes);
        atomicAdd(&a.stats[4], (unsigned long long)c.tris);
        atomicAdd(&a.stats[5], (unsigned long long)c.spills);
        atomicAdd(&a.stats[7], (unsigned long long)c.any_nodes);
        atomicAdd(&a.stats[8], (unsigned long long)c.any_tris);
    }
}

// One ray per thread, no refill: a grid of blocks over every slot the queues
// can hold (blocks past the queue size exit at once). For passes of few
// paths, where a persistent launch's time is mostly its lanes' tails.
template <int STACK, bool MULTI>
__global__ void __launch_bounds__(kBlock) k_trace_flat(RenderArgs a, int b) {
    __shared__ unsigned long long stk[STACK * kBlock];
    const int cur = b & 1;
    const uint32_t ns = a.counts[C_SH + (cur ^ 1)];
    const uint32_t ne = a.counts[C_EXT + cur];
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        a.counts[C_EXT + (cur ^ 1)] = 0u;
        a.counts[C_SH + cur] = 0u;
        a.counts[C_FETCH + (cur ^ 1)] = 0u;
        atomicAdd(&a.stats[0], (unsigned long long)ne);
        atomicAdd(&a.stats[1], (unsigned long long)ns);
    }
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= ns + ne) return;
    LaneStack ls;
    ls.lds = (lds_u64 *)(stk + threadIdx.x);
    ls.lds_stride = kBlock;
    ls.cap = STACK;
    ls.spill = a.spill + i;
    ls.spill_stride = a.spill_stride;
    TravCounters c;
    TraceSrc<MULTI> src{a.sq_o,    a.sq_d, a.sq_tgt, a.pend,    a.rq_o[cur], a.rq_d[cur], a.rq_sid[cur],
                        a.hits,    ns,     a.cap,    a.items,   a.st_item,   a.n_items,   a.sc, a.sq_occ, a.hit_uv};
    v3 o, d;
    float tmin, tmax;
    uint32_t anyw;
    TraceScene sc;
    src.load(i, o, d, tmin, tmax, anyw, sc);
    const bool any = (int32_t)anyw >= 0;
    TravState s;
    trav_init(s, o, d, tmin, tmax);
    while (!trav_step<false>(sc, s, ls, 1 << 30, c, any, any)) {
    }
    src.store(i, s);
}

// traversal modes of a pass (RenderArgs::trav_mode)
constexpr int kTravBudget = 0;     // persistent, refill after trav_budget iterations, static chunks
constexpr int kTravPrefetch = 1;   // persistent, each lane's next ray loaded ahead
constexpr int kTravFlat = 2;       // one ray per thread (counter passes use kTravPrefetch)

#ifndef FL_SHADE_TU
template <int S, int WV, bool M>
void launch_trace_m(bool count, dim3 g, hipStream_t st, const RenderArgs &a, int j) {
    if (a.trav_mode == kTravFlat && !count) {
        hipLaunchKernelGGL((k_trace_flat<S, M>), dim3((unsigned)((3ull * a.cap + kBlock - 1) / kBlock)), dim3(kBlock), 0,
                           st, a, j);
    } else if (a.trav_mode != kTravBudget) {
        if (count) hipLaunchKernelGGL((k_trace<true, S, WV, true, M>), g, dim3(kBlock), 0, st, a, j);
        else hipLaunchKernelGGL((k_trace<false, S, WV, true, M>), g, dim3(kBlock), 0, st, a, j);
    } else {
        if (count) hipLaunchKernelGGL((k_trace<true, S, WV, false, M>), g, dim3(kBlock), 0, st, a, j);
        else hipLaunchKernelGGL((k_trace<false, S, WV, false, M>), g, dim3(kBlock), 0, st, a, j);
    }
}
// batches (several items) run the MULTI instantiations
template <int S, int WV>
void launch_trace_t(bool count, dim3 g, hipStream_t st, const RenderArgs &a, int j) {
    if (a.n_items > 1) launch_trace_m<S, WV, true>(count, g, st, a, j);
    else launch_trace_m<S, WV, false>(count, g, st, a, j);
}
template <int S, int WV>
const void *trace_fn(bool prefetch, bool multi) {
    if (multi)
        return prefetch ? reinterpret_cast<const void *>(k_trace<false, S, WV, true, true>)
                        : reinterpret_cast<const void *>(k_trace<false, S, WV, false, true>);
    return prefetch ? reinterpret_cast<const void *>(k_trace<false, S, WV, true, false>)
                    : reinterpret_cast<const void *>(k_trace<false, S, WV, false, false>);
}

// traversal variants (LDS stack entries, minimum waves per SIMD); the first
// is the default, the rest are tuning choices
struct TraceVariant {
    int stack, waves;
    void (*launch)(bool, dim3, hipStream_t, const RenderArgs &, int);
    const void *(*fn)(bool, bool);
};
const TraceVariant kTraceVariants[] = {
    {8, 5, launch_trace_t<8, 5>, trace_fn<8, 5>},    {12, 1, launch_trace_t<12, 1>, trace_fn<12, 1>},
    {8, 6, launch_trace_t<8, 6>, trace_fn<8, 6>},    {8, 8, launch_trace_t<8, 8>, trace_fn<8, 8>},
    {6, 8, launch_trace_t<6, 8>, trace_fn<6, 8>},    {16, 1, launch_trace_t<16, 1>, trace_fn<16, 1>},
};
#endif  // !FL_SHADE_TU

// ---------------------------------------------------------------- guide ----
// Hardware exp, reciprocal and reciprocal square root (about 1 ulp) in the
// lobes' decode and density: the estimator needs the density to be that of
// the sampling, which holds to fp32 rounding whichever rounding the
// transcendental functions have.
__device__ __forceinline__ float hw_rcp(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ float hw_rsq(float x) { return __builtin_amdgcn_rsqf(x); }

// The lobes of a guided vertex live in its wave's LDS region for the rest of
// that vertex's shading (kLobeRows rows of 64 lanes; `lb` points at the
// lane's column). Row 6 l + j of lobe l: j = 0..2 mean direction, 3 kappa,
// 4 cumulative weight (of lobes 0..l), 5 weight times the vMF normalisation.
// Sampling picks a lobe by the cumulative weights and the density sums the
// scaled lobes, so both use the same mixture to fp32 rounding.
__device__ __forceinline__ float lobe_val(const float *lb, int l, int j) { return lb[(kLobeVals * l + j) * 64]; }

__device__ __forceinline__ float guide_pdf_at(const float *lb, v3 w) {
    float p = 0.f;
    // one lobe per iteration: 5 values live instead of 40 in code that is
    // already short of registers
#pragma unroll 1
    for (int l = 0; l < kGLobes; ++l) {
        const v3 mu(lobe_val(lb, l, 0), lobe_val(lb, l, 1), lobe_val(lb, l, 2));
        p += lobe_val(lb, l, 5) * __expf(lobe_val(lb, l, 3) * (dot(mu, w) - 1.f));
    }
    return p;
}

__device__ __forceinline__ v3 guide_sample_at(const float *lb, float u_lobe, float u1, float u2) {
    int sel = kGLobes - 1;
    for (int l = 0; l < kGLobes - 1; ++l)
        if (u_lobe < lobe_val(lb, l, 4)) { sel = l; break; }
    VmfLobe L;
    L.mu = v3(lobe_val(lb, sel, 0), lobe_val(lb, sel, 1), lobe_val(lb, sel, 2));
    L.kappa = lobe_val(lb, sel, 3);
    L.w = 1.f;
    return vmf_sample(L, u1, u2);
}

// Materials the guide serves: the non-specular ones.
__device__ __forceinline__ bool guidable(int kind) { return kind == MAT_LAMBERT || kind == MAT_CONDUCTOR; }

// Network inputs of a shading point: position in the unit scene box, the
// geometric normal turned toward the incoming ray, hero wavelength in the
// spectral range, roughness (GGX alpha; 1 for Lambert).
__device__ __forceinline__ void guide_features(const RenderArgs &a, v3 p, v3 ng, v3 d, float lam0, const DevMaterial &m,
                                               float pos01[3], v3 &n_in, float &lam01, float &rough) {
    const float pp[3] = {p.x, p.y, p.z};
    for (int k = 0; k < 3; ++k) pos01[k] = fminf(fmaxf((pp[k] - a.box_lo[k]) * a.box_inv[k], 0.f), 1.f);
    n_in = dot(ng, d) < 0.f ? ng : -ng;
    lam01 = (lam0 - a.grid.lo) / (a.grid.hi - a.grid.lo);
    rough = m.kind == MAT_CONDUCTOR ? m.ax : 1.f;
}

// The network for the 64 vertices of a wave, inside the guided shading
// kernel: the encoded inputs go into the wave's LDS region (bf16 rows), the
// four hidden layers run on the matrix cores in place with the bf16 weights
// read through the caches (41 KB, shared by every wave of the device), the
// head's 40 outputs go to the same region as fp32 rows, and each lane
// decodes its own row into its lobe column. Lanes without a query feed zeros
// and get no lobes. All lanes of the wave call this; `reg` is the wave's
// region (kGuideRegion floats).
constexpr int kGuideRaw = 41;   // fp32 row stride of the head outputs (odd: no bank conflicts)
constexpr int kGuideRegion = kLobeRows * 64;   // floats per wave: the largest of the three uses
static_assert(kGuideRegion * 4 >= 64 * kGRow * 2 && kGuideRegion >= 64 * kGuideRaw, "the region holds every stage");
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}
// W: the five bf16 weight matrices at row stride ws (LDS or global), matrix
// k at W + k wm; B: the five bias vectors, k at B + k bm.
__device__ __forceinline__ void guide_wave_mlp(float *reg, bool want, const float pos01[3], v3 n_in, float lam01,
                                               float rough, const __bf16 *W, int ws, int wm, const float *B, int bm) {
    const int q = threadIdx.x & 63;
    __bf16 *sT = reinterpret_cast<__bf16 *>(reg);
    // The weights and biases are the same for every vertex: without this
    // empty asm, which the compiler must assume changes the offset, it hoists
    // their loads (some 150 values per lane) out of the shading loop and keeps
    // them live across all of it.
    int off = 0;
    asm volatile("" : "+s"(off));
    const __bf16 *gW = W + off;
    const float *gP = B + off;
    encode_row_bf16(sT, q, want, pos01, n_in, lam01, rough);
    wave_sync();
#pragma unroll 1
    for (int k = 0; k < 4; ++k) {
        layer_fwd_inreg(gW + k * wm, ws, gP + k * bm, sT);
        wave_sync();
    }
    {   // head: rows 0..47 of the last matrix (40 used), into fp32 rows that
        // overlap the tile: the tile's fragments are all in registers first,
        // then each 16-row tile of outputs is computed and stored
        const int l = threadIdx.x & 63, r16 = l & 15, kq = (l >> 4) * 8;
        bf16x8 hq[2][4];
#pragma unroll
        for (int ks = 0; ks < 2; ++ks)
#pragma unroll
            for (int nt = 0; nt < 4; ++nt)
                hq[ks][nt] = *reinterpret_cast<const bf16x8 *>(sT + (nt * 16 + r16) * kGRow + ks * 32 + kq);
        const __bf16 *W4 = gW + 4 * wm;
        const float *bh = gP + 4 * bm;
#pragma unroll
        for (int mt = 0; mt < 3; ++mt) {
            f32x4 acc[4];
#pragma unroll
            for (int nt = 0; nt < 4; ++nt) acc[nt] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int ks = 0; ks < 2; ++ks) {
                const bf16x8 wa = *reinterpret_cast<const bf16x8 *>(W4 + (mt * 16 + r16) * ws + ks * 32 + kq);
#pragma unroll
                for (int nt = 0; nt < 4; ++nt)
                    acc[nt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wa, hq[ks][nt], acc[nt], 0, 0, 0);
            }
#pragma unroll
            for (int nt = 0; nt < 4; ++nt)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int o = acc_m(mt, r);
                    if (o < kGOut) reg[acc_n(nt) * kGuideRaw + o] = acc[nt][r] + bh[o];
                }
        }
    }
    wave_sync();
    // decode (guide_decode's mapping): mu = m / |m|, kappa = exp(clamp(raw)),
    // weights the softmax of the logits; all of the lane's row is read before
    // any lane writes its lobe column over the rows
    float rv[kGOut];
    if (want)
#pragma unroll
        for (int o = 0; o < kGOut; ++o) rv[o] = reg[q * kGuideRaw + o];
    wave_sync();
    if (want) {
        float mx = -3.0e38f;
#pragma unroll
        for (int l = 0; l < kGLobes; ++l) mx = fmaxf(mx, rv[5 * l + 4]);
        float sum = 0.f;
#pragma unroll
        for (int l = 0; l < kGLobes; ++l) {
            rv[5 * l + 4] = __expf(rv[5 * l + 4] - mx);
            sum += rv[5 * l + 4];
        }
        const float inv_sum = hw_rcp(sum);
        float cum = 0.f;
#pragma unroll
        for (int l = 0; l < kGLobes; ++l) {
            const float *r = rv + 5 * l;
            const float len2 = r[0] * r[0] + r[1] * r[1] + r[2] * r[2];
            const bool ok = len2 > 1e-12f;
            const float il = ok ? hw_rsq(len2) : 0.f;
            const float kappa = __expf(fminf(fmaxf(r[3], kGKappaLo), kGKappaHi));
            const float w = r[4] * inv_sum;
            cum += w;
            // vMF normalisation kappa / (2 pi (1 - exp(-2 kappa))), kappa >= exp(-3)
            const float norm = kappa * hw_rcp(2.f * kPi * (1.f - __expf(-2.f * kappa)));
            float *c = reg + kLobeVals * l * 64 + q;
            c[0] = ok ? r[0] * il : 0.f;
            c[64] = ok ? r[1] * il : 0.f;
            c[128] = ok ? r[2] * il : 1.f;
            c[192] = kappa;
            c[256] = cum;
            c[320] = w * norm;
        }
    }
}

// Online training runs in R passes of doubling size (R = the guide's
// train_rounds, 3 by default) over the first train_spp samples per pixel;
// pass r ends after train_spp (2^(r+1) - 1) / (2^R - 1) samples
// (r = R - 1 ends at train_spp).
inline uint32_t train_round_end(uint32_t train_spp, uint32_t rounds, int r) {
    if (r < 0) return 0u;
    const uint64_t den = (1ull << rounds) - 1ull, num = (2ull << r) - 1ull;
    return (uint32_t)std::max<uint64_t>((uint64_t)r + 1u, (uint64_t)train_spp * num / den);
}

// Training record weights after a pass: the radiance the path gathered beyond
// the vertex, plus the light-sampled share of the emission its continuation
// hit, hero lane = beta f cos L_i / p_sample with the full incident radiance.
// The throughput beta before the vertex is kept in the weight: a vertex's
// share of the loss is its share of the pixel's value, so vertices the image
// hardly sees (deep in a path) do not steer the network. Records whose
// continuation was never shaded keep weight 0.
__global__ void k_finish_records(RenderArgs a, uint32_t n) {
    const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    float *R = a.recs + (size_t)r * 16;
    const uint32_t sid = __float_as_uint(R[13]);
    const float Lf = a.st_L[sid].x + a.pend[sid].x + a.pend[a.cap + sid].x;
    const float Lb = R[12];
    float w = 0.f;
    if (Lb >= 0.f) w = fmaxf(Lf - Lb, 0.f) + R[15];
    R[14] = (w < 3.0e38f) ? w : 0.f;
}

// ---------------------------------------------------------------- shade ----
__device__ __forceinline__ void setup_bsdf(const RenderArgs &a, const float *spec, const DevMaterial &m,
                                           const Lanes &ln, Bsdf &b) {
    b.kind = m.kind;
    b.ggx.ax = m.ax;
    b.ggx.ay = m.ay;
    b.weight = m.weight;
    b.vndf_pdf = (a.flags & FL_RENDER_VNDF_PDF) != 0;
    const int S = a.grid.stride;
    const float *ta = spec + (size_t)m.spec_a * S;
    const float *tb = spec + (size_t)m.spec_b * S;
    for (int k = 0; k < 4; ++k) {
        b.s1[k] = spec_at(ta, ln, k);
        b.s2[k] = m.kind == MAT_LAMBERT ? 0.f : spec_at(tb, ln, k);
    }
    // dielectric: Cauchy n(lambda) = A + B / lambda^2 [um] at the hero wavelength
    float lum = ln.lam[0] * 1e-3f;
    b.ior = m.cauchy_b != 0.f ? m.eta_bsdf + m.cauchy_b / (lum * lum) : m.eta_bsdf;
    b.dispersive = m.cauchy_b != 0.f;
}

// A standard material's spectra at one vertex, read from the scene's tables
// when a lobe needs them (they are not all kept in registers).
struct TabSpec {
    const float *spec;
    const uint32_t *slot;
    int gs;
    Lanes ln;
    bool parity;
    const uint32_t *tex;     // the material's texture per spectrum (StdParams::tex); null: none
    const float *tv;         // this thread's textured lane values in LDS: [which * 4 + k] at stride kBlock
    __device__ float at(int which, int k) const {
        if (tex && which < SS_ETA && tex[which]) return tv[(which * 4 + k) * kBlock];
        return spec_at(spec + (size_t)slot[which] * gs, ln, k);
    }
    // ThinFilmInteference samples 60 wavelengths 400, 405, ... (thinfilm.cpp:38): in
    // parity mode a lane sees its bin's; otherwise its own wavelength
    __device__ float tf_lambda(int k) const { return parity ? 400.f + 5.f * (float)ln.idx[k] : ln.lam[k]; }
};

constexpr uint32_t kNonSpecular = (BSDF_REFLECTION | BSDF_TRANSMISSION | BSDF_DIFFUSE | BSDF_GLOSSY);
constexpr uint32_t kAllLobes = kNonSpecular | BSDF_SPECULAR;

// BSDF::evaluate / pdf with flags BSDF_ALL & ~BSDF_SPECULAR for either kind of material
__device__ __forceinline__ void shade_eval_nonspec(const Bsdf &b, const StdParams *sd, const TabSpec &tsp, v3 wo, v3 wi,
                                                   bool refl, float f[4]) {
    if (sd) std_bsdf_eval(*sd, tsp, b.vndf_pdf, wo, wi, refl, kNonSpecular, f);
    else bsdf_eval_nonspec(b, wo, wi, refl, f);
}

__device__ __forceinline__ float shade_pdf_nonspec(const Bsdf &b, const StdParams *sd, v3 wo, v3 wi) {
    return sd ? std_bsdf_pdf(*sd, b.vndf_pdf, wo, wi, kNonSpecular) : bsdf_pdf_nonspec(b, wo, wi);
}

// The value of light Lj toward the shading point for delta lights
// (sample_li of point.cpp:12-20, spot.cpp:23-31, directional.cpp:15-24):
// the direction wi toward the light, the unnormalized segment to its
// visibility point and the factor intensity * falloff / decay that scales
// the light's color. Returns false when the light gives nothing there.
__device__ __forceinline__ bool delta_light_sample(const DevLight &Lj, v3 p, v3 &wi, v3 &seg, float &scale) {
    if (Lj.type == FL_LIGHT_DIRECTIONAL) {
        wi = ld3(Lj.p0);
        seg = wi * (2.f * 1000.f);          // pOutside = p + dirIllum * 2 worldRadius
        scale = Lj.intensity;
        return true;
    }
    const v3 pos = ld3(Lj.p0);
    seg = pos - p;
    if (length2(seg) == 0.f) return false;
    wi = normalize(seg);
    const float dl = length(seg);
    float s_ = Lj.intensity / (dl * dl);    // evaluate_decay_rate: d^2 whatever the decay rate
    if (Lj.type == FL_LIGHT_SPOT) {
        // SpotIlluminant::Falloff(-wi), spot.cpp:41-55
        const float ct = dot(normalize(-wi), normalize(ld3(Lj.p1)));
        float fo;
        if (ct < Lj.cos_cone) fo = 0.f;
        else if (ct > Lj.cos_penumbra) fo = 1.f;
        else {
            const float t = fminf(fmaxf((ct - Lj.cos_cone) / (Lj.cos_penumbra - Lj.cos_cone), 0.f), 1.f);
            fo = t * t * (3.f - 2.f * t);
        }
        s_ = Lj.intensity * fo / (dl * dl);
    }
    scale = s_;
    return scale != 0.f;
}

// One NEE sample of light `li` (estimate_direct's light half, integrator.cpp:29-59):
// returns true with the (unoccluded) contribution, the contribution if the
// segment turns out occluded (Li = shadowColor) and the shadow segment.
__device__ __forceinline__ bool nee_sample(const RenderArgs &a, const SceneRef &S, const Bsdf &b, const StdParams *sd,
                                           const TabSpec &tsp, const Lanes &ln, uint32_t li, float sel,
                                           float ux, float uy, v3 p, v3 perr, v3 ng, v3 fs, v3 ft, v3 woW, v3 wo,
                                           const float beta[4], float contrib[4], float occ[4], float4 &sh_o,
                                           float4 &sh_d, const float *lb, float mix) {
    const DevLight &Lj = S.lights[li];
    const float *te = S.spec + (size_t)Lj.em_slot * a.grid.stride;
    const float *ts = Lj.shadow_slot ? S.spec + (size_t)Lj.shadow_slot * a.grid.stride : nullptr;
    if (Lj.type != 0u) {
        // delta light: pdf 1, no MIS (integrator.cpp:53-54)
        v3 wi, seg;
        float scale;
        if (!delta_light_sample(Lj, p, wi, seg, scale)) return false;
        v3 wil(dot(wi, fs), dot(wi, ft), dot(wi, ng));
        bool refl = dot(wi, ng) * dot(woW, ng) > 0.f;
        float f[4];
        shade_eval_nonspec(b, sd, tsp, wo, wil, refl, f);
        const float cw = absdot(wi, ng);
        bool black = true;
        for (int k = 0; k < 4; ++k) { f[k] = f[k] * cw; black = black && f[k] == 0.f; }
        if (black) return false;
        for (int k = 0; k < 4; ++k) {
            contrib[k] = beta[k] * f[k] * (spec_at(te, ln, k) * scale) / sel;
            occ[k] = ts ? beta[k] * f[k] * spec_at(ts, ln, k) / sel : 0.f;
        }
        // kindle_ray_to(target): d = target - p, origin offset along d
        const v3 sd = Lj.type == FL_LIGHT_DIRECTIONAL ? (p + seg) - p : seg;
        v3 so = offset_point_origin(p, perr, ng, sd);
        sh_o = make_float4(so.x, so.y, so.z, kShadowEps);
        sh_d = make_float4(sd.x, sd.y, sd.z, 1.f - kShadowEps);
        return true;
    }
    // Triangle::sample + Shape::sample(ref), triangle.cpp:134-154, shape.cpp:28-41
    float uu = sqrtf(ux);
    float u1 = 1.f - uu, u2 = uy * uu;
    v3 lp0 = ld3(Lj.p0), lp1 = ld3(Lj.p1), lp2 = ld3(Lj.p2), lnn = ld3(Lj.n);
    v3 ps = lp0 * u1 + lp1 * u2 + lp2 * (1.f - u1 - u2);
    v3 wun = ps - p;
    if (length2(wun) == 0.f) return false;
    v3 wi = normalize(wun);
    float pdf_l = (1.f / Lj.area) * (length2(p - ps) / absdot(lnn, -wi));
    if (isinf(pdf_l)) pdf_l = 0.f;
    pdf_l *= sel;
    if (!(pdf_l > 0.f)) return false;
    if (!(dot(lnn, -wi) > 0.f)) return false;                       // DiffuseAreaIlluminant::L
    v3 wil(dot(wi, fs), dot(wi, ft), dot(wi, ng));
    bool refl = dot(wi, ng) * dot(woW, ng) > 0.f;
    float f[4];
    shade_eval_nonspec(b, sd, tsp, wo, wil, refl, f);
    float cw = absdot(wi, ng);
    float pb = shade_pdf_nonspec(b, sd, wo, wil);
    // guided vertices sample the mixture (1 - mix) BSDF + mix guide: that is
    // the density the light sample's MIS weight must compete with
    if (lb) pb = (1.f - mix) * pb + mix * guide_pdf_at(lb, wi);
    bool black = true;
    for (int k = 0; k < 4; ++k) { f[k] = f[k] * cw; black = black && f[k] == 0.f; }
    if (black) return false;
    float w = power_heuristic(pdf_l, pb);
    for (int k = 0; k < 4; ++k) {
        contrib[k] = beta[k] * f[k] * (spec_at(te, ln, k) * Lj.intensity) * w / pdf_l;
        occ[k] = ts ? beta[k] * f[k] * spec_at(ts, ln, k) * w / pdf_l : 0.f;
    }
    v3 so = offset_point_origin(p, perr, ng, wun);                 // kindle_ray_to
    sh_o = make_float4(so.x, so.y, so.z, kShadowEps);
    sh_d = make_float4(wun.x, wun.y, wun.z, 1.f - kShadowEps);
    return true;
}

// NormalIntegrator::Li (normalintegrator.cpp:11-26): from_reflectance_rgb of
// the hit's normal (spectrumoperations.cpp:113-147) at the path's four
// wavelengths. basis: the 7 basis curves, GS floats apart.
__device__ __forceinline__ void normal_spectrum(const float *basis, int GS, v3 n, const Lanes &ln, float L[4]) {
    for (int k = 0; k < 4; ++k) L[k] = rgb_spectrum_lane(basis, GS, n.x, n.y, n.z, ln, k);
}

// WAVES: minimum waves per SIMD requested from the register allocator (2 lets
// it use all the registers it wants; 4 caps at 128 VGPRs with a few spills).
// Shading sort (FL_RENDER_SORT_SHADING): each block orders its queue entries
// by material kind (counting sort in LDS) before shading them, so a wave runs
// one BSDF's code instead of every kind present in its 64 entries. Per-path
// results do not depend on the order; only the order of the next queue
// changes. It is off by default: its key costs three dependent loads (hit,
// triangle, material) per entry and three barriers per block step, and the
// shading was faster without it on every scene measured (spheres 20.06 ->
// 19.60 ms, the room 50.1 -> 44.1 ms, Cornell, the standard-surface box, the
// terrain and guided shading 1-5%).
constexpr uint32_t kSortKeys = 6;   // miss / no material, Lambert, conductor, dielectric, standard, past the end

__device__ __forceinline__ uint32_t shade_key(const RenderArgs &a, int cur, uint32_t i, uint32_t n) {
    if (i >= n) return kSortKeys - 1;
    const uint32_t rec = __float_as_uint(a.hits[i].y);
    if (rec == kNoHit) return 0;
    const SceneRef S = scene_of(a, item_of(a, a.rq_sid[cur][i]));
    const uint32_t mat = reinterpret_cast<const uint4 *>(S.sc.tris + rec)[1].w;
    if (mat == kNoMaterial) return 0;
    return (uint32_t)S.mats[mat].kind;   // MAT_LAMBERT 1, MAT_CONDUCTOR 2, MAT_DIELECTRIC 3
}

// A path's pixel (y W + x) and sample index, from its id: k_generate lays
// out an item's paths by sample layer of npix padded pixels in 8x8 tiles.
__device__ __forceinline__ uint32_t path_pixel(const RenderArgs &a, uint32_t sid) {
    const ItemDev &it = a.items[item_of(a, sid)];
    const uint32_t local = sid - it.path_base, pl = local % it.npix;
    const uint32_t tile = pl >> 6, within = pl & 63u;
    const uint32_t x = (tile % it.tiles_x) * 8u + (within & 7u), y = (tile / it.tiles_x) * 8u + (within >> 3);
    return y * it.cam.W + x;
}
__device__ __forceinline__ uint32_t path_sample(const RenderArgs &a, uint32_t sid) {
    const ItemDev &it = a.items[item_of(a, sid)];
    return it.sample_base + (sid - it.path_base) / it.npix;
}

// GUIDED: the render has a guide (mixture sampling at guided vertices and
// training records); other renders run a kernel without that code.
// STD: the render has standard-surface materials or image textures. Their
// code is only compiled into that variant, so renders without them keep the
// smaller kernel (registers and instruction cache).
// Threads per block of a shading instantiation: guided renders without
// standard surfaces run 8-wave blocks, one per CU, so that one LDS copy of the
// network's weights serves 8 waves (145 KB of LDS with their regions).
constexpr int kGuideBlock = 512;
constexpr int shade_block(bool std_, bool guided) { return guided && !std_ ? kGuideBlock : kBlock; }

#ifdef FL_SHADE_TU
template <int WAVES, bool STD, bool GUIDED>
__global__ void __launch_bounds__(shade_block(STD, GUIDED)) __attribute__((amdgpu_waves_per_eu(WAVES, 8)))
k_shade(RenderArgs a, int b) {
    constexpr int BS = shade_block(STD, GUIDED);
    constexpr bool LDS_W = GUIDED && !STD;   // the network's weights in LDS
    __shared__ uint32_t s_cnt[kSortKeys];
    __shared__ uint16_t s_perm[BS];
    // per-block queue reservations: [0] shadow rays, [1] continuations,
    // [2] training records
    __shared__ uint32_t s_app[3][BS / 64 + 1];
    // a textured standard material's kd, ks, kt, sheen and coat at the vertex's
    // four wavelengths, one column per thread (20 KB: registers would hold
    // them across the whole shading of every vertex)
    __shared__ float s_tv[STD ? SS_ETA * 4 * BS : 1];
    // guided renders: each wave's region for the network and its lobes (12 KB)
    __shared__ __attribute__((aligned(16))) float s_guide[GUIDED ? (BS / 64) * kGuideRegion : 1];
    // and, without standard surfaces, the bf16 weights (rows padded to kGRow
    // against bank conflicts) and the biases, loaded once per block
    __shared__ __attribute__((aligned(16))) __bf16 s_gw[LDS_W ? kGMats * 64 * kGRow : 8];
    __shared__ float s_gb[LDS_W ? kGMats * 64 : 1];
    unsigned long long n_queries = 0;   // guided vertices of this wave (counted by lane 0)
    if (LDS_W && a.guided) {
        for (int e = threadIdx.x; e < kGMats * 64 * 8; e += BS) {   // 16-byte pieces of W[k][o][0..63]
            const int k = e / 512, o = (e / 8) % 64, c = e % 8;
            *reinterpret_cast<bf16x8 *>(&s_gw[(k * 64 + o) * kGRow + 8 * c]) =
                *reinterpret_cast<const bf16x8 *>(a.gW + k * 4096 + o * 64 + 8 * c);
        }
        for (int e = threadIdx.x; e < kGMats * 64; e += BS) s_gb[e] = a.gP[(e / 64) * kGLayerParams + 4096 + e % 64];
        __syncthreads();
    }
    const uint32_t stride = gridDim.x * BS;
    const int cur = b & 1, nxt = cur ^ 1;
    const uint32_t n = a.counts[C_EXT + cur];
    const uint32_t nround = (n + BS - 1) / BS * BS;
    const int GS = a.grid.stride;
    const bool sort = (a.flags & FL_RENDER_SORT_SHADING) != 0;
    for (uint32_t base = blockIdx.x * BS; base < nround; base += stride) {
        uint32_t i = base + threadIdx.x;
        if (sort) {
            const uint32_t key = shade_key(a, cur, i, n);
            if (threadIdx.x < kSortKeys) s_cnt[threadIdx.x] = 0u;
            __syncthreads();
            const uint32_t rank = atomicAdd(&s_cnt[key], 1u);
            __syncthreads();
            uint32_t off = 0;
            for (uint32_t k = 0; k < key; ++k) off += s_cnt[k];
            s_perm[off + rank] = (uint16_t)threadIdx.x;
            __syncthreads();
            i = base + s_perm[threadIdx.x];
            __syncthreads();   // s_cnt and s_perm are rewritten by the next iteration
        }
        const bool active = i < n;
        const SceneRef S = scene_of(a, active ? item_of(a, a.rq_sid[cur][i]) : 0u);
        // -------- phase 1a: the hit's geometry and the path's wavelength
        bool cont = false, mat_hit = false, rec_want = false, hit = false, fresh = false;
        uint32_t rec_prev = 0;   // index + 1 of the previous vertex's training record
        float beta0 = 0.f;
        uint32_t sid = 0, bounce = 0, flags = 0, rb = 0, mkind = 0, light = 0;
        float beta[4] = {0.f, 0.f, 0.f, 0.f}, L[4] = {0.f, 0.f, 0.f, 0.f};
        float lam0 = 0.f, eta_scale = 1.f, prev_pdf = 0.f, eta_bsdf = 1.f, t = 0.f;
        v3 p, ng, d, no, nd, perr;
        float2 uv = make_float2(0.f, 0.f);   // texture coordinates of the hit (textured scenes)
        Lanes ln;
        if (active) {
            const float4 ro = a.rq_o[cur][i];
            const Packed3 rd = a.rq_d[cur][i];
            sid = a.rq_sid[cur][i];
            const float2 hh = a.hits[i];
            const float4 misc = a.st_misc[sid];
            lam0 = misc.x; eta_scale = misc.y; prev_pdf = misc.z;
            const uint32_t bits = __float_as_uint(misc.w);
            bounce = bits & 0xffu;
            flags = bits & ~0xffu;
            fresh = (flags & PF_FRESH) != 0u;
            flags &= ~PF_FRESH;
            const uint32_t rec = __float_as_uint(hh.y);
            t = hh.x;
            d = v3(rd.x, rd.y, rd.z);
            if (rec != kNoHit) {
                const uint4 *tp = reinterpret_cast<const uint4 *>(S.sc.tris + rec);
                uint4 q0 = tp[0], q1 = tp[1], q2 = tp[2];
                v3 p0(__uint_as_float(q0.x), __uint_as_float(q0.y), __uint_as_float(q0.z));
                v3 p1(__uint_as_float(q1.x), __uint_as_float(q1.y), __uint_as_float(q1.z));
                v3 p2(__uint_as_float(q2.x), __uint_as_float(q2.y), __uint_as_float(q2.z));
                mkind = q1.w;
                light = q2.w;
                p = v3(ro.x, ro.y, ro.z) + d * t;                  // ray.point_at_t(t)
                ng = normalize(cross(p1 - p0, p2 - p0));
                if (STD && S.tex) {
                    const float2 bc = a.hit_uv[i];
                    uv = hit_uv(*S.tex, q0.w, bc.x, bc.y);
                }
                hit = true;
                // a material vertex unless the first hit ends the path
                // (NormalIntegrator), the MIS ray of the last vertex is being
                // closed, or the primitive has no BSDF
                mat_hit = !(a.flags & FL_RENDER_NORMALS) && !(flags & PF_LAST) && mkind != kNoMaterial;
            }
        }
        // -------- the guide's lobes at this vertex (guided passes): the
        // network runs for the wave's guidable vertices together, before the
        // path state is loaded, so that little is live across it
        const float *lb = nullptr;
        if constexpr (GUIDED) {
            if (a.guided) {
                const bool want = mat_hit && bounce < a.guide_bounces && guidable(S.mats[mkind].kind);
                const unsigned long long wm = __ballot(want);
                if (wm) {
                    if (__lane_id() == 0) n_queries += (unsigned long long)__popcll(wm);
                    float pos01[3] = {0.f, 0.f, 0.f}, lam01 = 0.f, rough = 0.f;
                    v3 n_in(0.f, 0.f, 0.f);
                    if (want) guide_features(a, p, ng, d, lam0, S.mats[mkind], pos01, n_in, lam01, rough);
                    float *reg = s_guide + (threadIdx.x >> 6) * kGuideRegion;
                    if constexpr (LDS_W)
                        guide_wave_mlp(reg, want, pos01, n_in, lam01, rough, s_gw, kGRow, 64 * kGRow, s_gb, 64);
                    else
                        guide_wave_mlp(reg, want, pos01, n_in, lam01, rough, a.gW, kGH, 4096, a.gP + 4096,
                                       kGLayerParams);
                    if (want) lb = reg + (threadIdx.x & 63);
                }
            }
        }
        // -------- phase 1b: path state, emission, pass-through
        if (active && fresh) {
            beta[0] = beta[1] = beta[2] = beta[3] = 1.f;
            L[0] = L[1] = L[2] = L[3] = 0.f;
        } else if (active) {
            const float4 beta4 = a.st_beta[sid], L4 = a.st_L[sid];
            const float4 pe0 = a.pend[sid], pe1 = a.pend[a.cap + sid];
            beta[0] = beta4.x; beta[1] = beta4.y; beta[2] = beta4.z; beta[3] = beta4.w;
            L[0] = L4.x + pe0.x + pe1.x; L[1] = L4.y + pe0.y + pe1.y;
            L[2] = L4.z + pe0.z + pe1.z; L[3] = L4.w + pe0.w + pe1.w;
        }
        if (active) {
            if (a.st_rec) {
                // the previous vertex's training record: radiance gathered up to
                // and including that vertex's own light samples
                rec_prev = a.st_rec[sid];
                if (rec_prev) {
                    a.recs[(size_t)(rec_prev - 1u) * 16 + 12] = L[0];
                    a.st_rec[sid] = 0u;
                }
            }
            if (hit) {
                ln = make_lanes(a.grid, lam0);
                // emission: L += beta * le(-dir) at bounce 0 / after specular; MIS otherwise
                if (a.flags & FL_RENDER_NORMALS) {
                    normal_spectrum(S.spec + (size_t)S.rgb_basis * GS, GS, ng, ln, L);
                    flags |= PF_LAST;   // NormalIntegrator: the first hit ends the path
                } else if (light) {
                    const DevLight &Lj = S.lights[light - 1];
                    if (dot(ng, -d) > 0.f) {
                        float w = 0.f;
                        if (bounce == 0 || (flags & PF_SPECULAR)) {
                            w = 1.f;
                        } else if (!(flags & PF_PASSED) && prev_pdf > 0.f) {
                            float dl = length(d);
                            float dist2 = (t * t) * (dl * dl);
                            float cosl = fabsf(dot(ng, d)) / dl;
                            float pl = dist2 / (cosl * Lj.area);
                            if (isinf(pl)) pl = 0.f;
                            if (!S.nee_all) pl = pl / (float)S.n_lights;
                            if (pl > 0.f) w = power_heuristic(prev_pdf, pl);
                        }
                        const float *te = S.spec + (size_t)Lj.em_slot * GS;
                        if (w > 0.f)
                            for (int k = 0; k < 4; ++k) L[k] += beta[k] * (spec_at(te, ln, k) * Lj.intensity) * w;
                        // the record of the previous vertex learns the full incident
                        // radiance: the share of this emission that light sampling
                        // covers under MIS is added to its target separately
                        if (rec_prev && w < 1.f)
                            a.recs[(size_t)(rec_prev - 1u) * 16 + 15] =
                                beta[0] * (spec_at(te, ln, 0) * Lj.intensity) * (1.f - w);
                    }
                }
                perr = vabs(p) * gamma_n(5);
                if (flags & PF_LAST) {
                    // the MIS ray of the last vertex: emission counted, path ends
                } else if (mkind == kNoMaterial) {
                    // no BSDF: isect.kindle_ray(ray.dir), bounces-- (pathintegrator.cpp:34-38)
                    no = offset_point_origin(p, perr, ng, d);
                    nd = d;
                    flags |= PF_PASSED;
                    cont = true;
                } else {
                    flags &= ~PF_PASSED;
                }
            }
        }
        // -------- phase 2: NEE, one sample per light (uniform_sample_all_lights)
        Bsdf bs_;
        v3 fs, ft, woW, wo;
        const StdParams *sd = nullptr;     // the lobe stack of a standard material
        if (mat_hit) {
            const DevMaterial m = S.mats[mkind];
            setup_bsdf(a, S.spec, m, ln, bs_);
            if (STD && m.kind == MAT_STANDARD) sd = S.stds + m.spec_a;
            if (STD && S.tex) {
                // ImageMapTexture::evaluate for each textured spectrum of the
                // material (lambert.cpp:12-13, standardsurface.cpp:84-85,115,146,155)
                const float *basis = S.spec + (size_t)S.rgb_basis * GS;
                if (m.tex) {
                    float rgb[3];
                    texel_rgb(*S.tex, m.tex - 1u, uv, rgb);
                    for (int k = 0; k < 4; ++k) bs_.s1[k] = rgb_spectrum_lane(basis, GS, rgb[0], rgb[1], rgb[2], ln, k);
                }
                if (sd)
                    for (int w = 0; w < SS_ETA; ++w)
                        if (sd->tex[w]) {
                            float rgb[3];
                            texel_rgb(*S.tex, sd->tex[w] - 1u, uv, rgb);
                            for (int k = 0; k < 4; ++k)
                                s_tv[(w * 4 + k) * BS + threadIdx.x] =
                                    rgb_spectrum_lane(basis, GS, rgb[0], rgb[1], rgb[2], ln, k);
                        }
            }
            eta_bsdf = m.eta_bsdf;
            coordinate_system(ng, &fs, &ft);                       // BSDF frame, bsdf.cpp:6-10 (ns = ng)
            fs = normalize(fs);
            woW = -d;
            wo = v3(dot(woW, fs), dot(woW, ft), dot(woW, ng));
            rb = rng_base(S.seed, path_pixel(a, sid), path_sample(a, sid), bounce);
        }
        const TabSpec tsp{S.spec, sd ? sd->slot : nullptr, GS, ln, a.grid.linear == 0,
                          (STD && sd && S.tex) ? sd->tex : nullptr, s_tv + (STD ? threadIdx.x : 0u)};
        const bool do_nee = mat_hit && S.n_lights > 0 &&
                            (sd ? std_num_components(*sd, kNonSpecular) > 0 : !is_specular_kind(bs_.kind));
        // both samples' shadow rays, appended together after the second
        float4 sh_ov[2], sh_dv[2], occv[2];
        bool okv[2], deltav[2];
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            float contrib[4] = {0.f, 0.f, 0.f, 0.f}, occ[4] = {0.f, 0.f, 0.f, 0.f};
            float4 sh_o = make_float4(0.f, 0.f, 0.f, 0.f), sh_d = sh_o;
            bool ok = false, nee_delta = false;
            const int nslots = S.nee_all ? (int)S.n_lights : 1;
            if (do_nee && j < nslots) {
                uint32_t li = (uint32_t)j;
                float sel = 1.f;
                if (!S.nee_all) {
                    li = min((uint32_t)(rng_u(rb, 7) * (float)S.n_lights), S.n_lights - 1u);
                    sel = 1.f / (float)S.n_lights;
                }
                ok = nee_sample(a, S, bs_, sd, tsp, ln, li, sel, rng_u(rb, 3 + 2 * j), rng_u(rb, 4 + 2 * j), p, perr, ng, fs, ft,
                                woW, wo, beta, contrib, occ, sh_o, sh_d, lb, a.guide_mix);
                nee_delta = S.lights[li].type != 0u;
            }
            if (active) {   // every live path rewrites both slots (0 when no sample)
                if (flags & PF_SECTERM) contrib[1] = contrib[2] = contrib[3] = occ[1] = occ[2] = occ[3] = 0.f;
                a.pend[(size_t)j * a.cap + sid] = make_float4(contrib[0], contrib[1], contrib[2], contrib[3]);
            }
            sh_ov[j] = sh_o;
            sh_dv[j] = sh_d;
            occv[j] = make_float4(occ[0], occ[1], occ[2], occ[3]);
            okv[j] = ok;
            deltav[j] = nee_delta;
        }
        {
            const unsigned long long m0 = __ballot(okv[0]), m1 = __ballot(okv[1]);
            const unsigned long long lt = (1ull << __lane_id()) - 1ull;
            const uint32_t c0 = (uint32_t)__popcll(m0);
            const uint32_t base = block_reserve<BS>(c0 + (uint32_t)__popcll(m1), &a.counts[C_SH + cur], s_app[0]);
#pragma unroll
            for (int j = 0; j < 2; ++j) {
                if (okv[j]) {
                    const uint32_t ss = base + (j ? c0 + (uint32_t)__popcll(m1 & lt) : (uint32_t)__popcll(m0 & lt));
                    a.sq_o[ss] = Packed3{sh_ov[j].x, sh_ov[j].y, sh_ov[j].z};
                    a.sq_d[ss] = Packed3{sh_dv[j].x, sh_dv[j].y, sh_dv[j].z};
                    a.sq_tgt[ss] = ((uint32_t)j * a.cap + sid) | (deltav[j] ? kVisClosest : 0u);
                    if (a.sq_occ) a.sq_occ[ss] = occv[j];
                }
            }
        }
        // -------- phase 3: BSDF sample (continuation and MIS BSDF ray), RR
        if (mat_hit) {
            BsdfSample bsm;
            bool ok;
            v3 wiW;
            beta0 = beta[0];
            float rr_guide = 1.f;
            if (lb) {
                // one-sample mixture: the guide with probability mix, else the
                // BSDF; the throughput divides by the mixture density at the
                // chosen direction, so the estimate stays unbiased wherever
                // the BSDF is non-zero
                bsm.pdf = 0.f;
                bsm.type = lobe_type(bs_.kind);
                bsm.secondary_terminated = false;
                v3 wl;
                bool have = wo.z != 0.f;
                if (have) {
                    if (rng_u(rb, 8) < a.guide_mix) {
                        wiW = guide_sample_at(lb, rng_u(rb, 9), rng_u(rb, 10), rng_u(rb, 11));
                        wl = v3(dot(wiW, fs), dot(wiW, ft), dot(wiW, ng));
                    } else {
                        BsdfSample bs2;
                        lobe_sample(bs_, wo, remap_u0(rng_u(rb, 0)), rng_u(rb, 1), bs2);
                        have = bs2.pdf != 0.f;
                        wl = bs2.wi;
                        wiW = fs * wl.x + ft * wl.y + ng * wl.z;
                    }
                }
                if (have) {
                    const bool refl = dot(wiW, ng) * dot(woW, ng) > 0.f;
                    bsdf_eval_nonspec(bs_, wo, wl, refl, bsm.f);
                    const float pb_dir = bsdf_pdf_nonspec(bs_, wo, wl);
                    bsm.pdf = (1.f - a.guide_mix) * pb_dir + a.guide_mix * guide_pdf_at(lb, wiW);
                    bsm.wi = wl;
                    // Russian roulette judges the throughput BSDF sampling would
                    // have given this direction (beta * f cos / p_bsdf), not the
                    // guided one: a direction the guide favours because much
                    // light arrives from it gets a small beta, and rouletting on
                    // that beta would end exactly the paths about to gather it
                    rr_guide = pb_dir > 0.f ? fminf(bsm.pdf / pb_dir, 1.0e4f) : 1.0e4f;
                }
                ok = bsm.pdf > 0.f;
                if (ok) {
                    bool black = true;
                    for (int k = 0; k < 4; ++k) black = black && bsm.f[k] == 0.f;
                    if (black) ok = false;
                }
            } else if (STD && sd) {
                // BSDF::sample over the lobe stack (component choice and u remap inside)
                bsm.wi = v3(0.f, 0.f, 0.f);
                std_bsdf_sample(*sd, tsp, bs_.vndf_pdf, wo, rng_u(rb, 0), rng_u(rb, 1), kAllLobes, bsm);
                ok = bsm.pdf != 0.f;
                if (ok) {
                    wiW = fs * bsm.wi.x + ft * bsm.wi.y + ng * bsm.wi.z;   // to_world
                    bool black = true;
                    for (int k = 0; k < 4; ++k) black = black && bsm.f[k] == 0.f;
                    if (black) ok = false;
                }
            } else {
                float u0 = remap_u0(rng_u(rb, 0)), u1 = rng_u(rb, 1);
                if (wo.z == 0.f) bsm.pdf = 0.f;
                else lobe_sample(bs_, wo, u0, u1, bsm);
                ok = bsm.pdf != 0.f;
                if (ok) {
                    wiW = fs * bsm.wi.x + ft * bsm.wi.y + ng * bsm.wi.z;   // to_world
                    if (!(bsm.type & BSDF_SPECULAR)) {
                        bool refl = dot(wiW, ng) * dot(woW, ng) > 0.f;
                        if (!refl) {
                            for (int k = 0; k < 4; ++k) bsm.f[k] = 0.f;
                        } else if (bs_.kind == MAT_CONDUCTOR) {
                            microfacet_f(bs_, wo, bsm.wi, bsm.f, &bsm.wh);
                        }
                    }
                    bool black = true;
                    for (int k = 0; k < 4; ++k) black = black && bsm.f[k] == 0.f;
                    if (black) ok = false;
                }
            }
            if (ok) {
                const uint32_t ty = bsm.type;
                bool brk = (a.depth_d == (int)bounce && (ty & BSDF_DIFFUSE)) ||
                           (a.depth_r == (int)bounce && (ty & (BSDF_SPECULAR | BSDF_GLOSSY))) ||
                           (a.depth_t == (int)bounce && (ty & BSDF_TRANSMISSION));
                if (brk && (ty & BSDF_SPECULAR)) ok = false;
                if (ok) {
                    if (brk) flags |= PF_LAST;
                    float cw = absdot(wiW, ng);
                    for (int k = 0; k < 4; ++k) beta[k] = beta[k] * bsm.f[k] * cw / bsm.pdf;
                    if (ty & BSDF_SPECULAR) flags |= PF_SPECULAR; else flags &= ~PF_SPECULAR;
                    if ((ty & BSDF_SPECULAR) && (ty & BSDF_TRANSMISSION))
                        eta_scale *= (dot(woW, ng) > 0.f) ? (eta_bsdf * eta_bsdf) : 1.f / (eta_bsdf * eta_bsdf);
                    if (lb) eta_scale *= rr_guide;   // eta_scale only scales the roulette's throughput
                    if (bsm.secondary_terminated) flags |= PF_SECTERM;
                    no = offset_point_origin(p, perr, ng, wiW);          // kindle_ray
                    nd = wiW;
                    prev_pdf = (ty & BSDF_SPECULAR) ? 0.f : bsm.pdf;
                    // A stack with specular lobes draws this non-specular
                    // direction with density S/(m+k) (k specular lobes among
                    // m+k), while the light sample's MIS weight competes with
                    // estimate_direct's BSDF density S/m (flags without
                    // BSDF_SPECULAR, integrator.cpp:63-66). The emission weight
                    // at the next hit takes the same S/m, so the two weights
                    // sum to one; beta already divides by the true density.
                    if (STD && sd && !(ty & BSDF_SPECULAR) &&
                        std_num_components(*sd, kAllLobes) != std_num_components(*sd, kNonSpecular))
                        prev_pdf = std_bsdf_pdf(*sd, bs_.vndf_pdf, wo, bsm.wi, kNonSpecular);
                    cont = true;
                    if (!brk) {
                        // Russian roulette, pathintegrator.cpp:77-83
                        float mx = beta[0] * eta_scale;
                        if (!(flags & PF_SECTERM))
                            for (int k = 1; k < 4; ++k) mx = fmaxf(mx, beta[k] * eta_scale);
                        if (mx < a.rr_threshold && (int)bounce > (int)a.rr_start) {
                            float q = fmaxf(0.05f, 1.f - mx);
                            if (rng_u(rb, 2) < q) cont = false;
                            else for (int k = 0; k < 4; ++k) beta[k] = beta[k] / (1.f - q);
                        }
                    }
                    bounce += 1;
                    rec_want = a.recs != nullptr && cont && guidable(bs_.kind);
                }
            }
        }
        if (active) {
            if (!cont) flags |= PF_DONE;
            if (flags & PF_SECTERM) { L[1] = L[2] = L[3] = 0.f; }
            a.st_L[sid] = make_float4(L[0], L[1], L[2], L[3]);
            a.st_misc[sid] = make_float4(lam0, eta_scale, prev_pdf, __uint_as_float(flags | bounce));
            if (cont) a.st_beta[sid] = make_float4(beta[0], beta[1], beta[2], beta[3]);
        }
        // -------- training record of this vertex (training passes only)
        if (GUIDED && a.recs) {
            const unsigned long long mr = __ballot(rec_want);
            const uint32_t ridx = block_reserve<BS>((uint32_t)__popcll(mr), a.rec_count, s_app[2]) +
                                  (uint32_t)__popcll(mr & ((1ull << __lane_id()) - 1ull));
            if (rec_want && ridx < a.rec_cap) {
                float *R = a.recs + (size_t)ridx * 16;
                float pos01[3], lam01, rough;
                v3 n_in;
                guide_features(a, p, ng, d, lam0, S.mats[mkind], pos01, n_in, lam01, rough);
                R[0] = pos01[0]; R[1] = pos01[1]; R[2] = pos01[2];
                R[3] = n_in.x; R[4] = n_in.y; R[5] = n_in.z;
                R[6] = lam01; R[7] = rough;
                R[8] = nd.x; R[9] = nd.y; R[10] = nd.z;
                R[11] = beta0;
                R[12] = -1.f;                      // radiance before: written at the next vertex
                R[13] = __uint_as_float(sid);
                R[14] = 0.f; R[15] = 0.f;
                a.st_rec[sid] = ridx + 1u;
            }
        }
        // -------- compaction of the continuation: one atomic per block
        const unsigned long long mc = __ballot(cont);
        const uint32_t slot = block_reserve<BS>((uint32_t)__popcll(mc), &a.counts[C_EXT + nxt], s_app[1]) +
                              (uint32_t)__popcll(mc & ((1ull << __lane_id()) - 1ull));
        if (cont) {
            a.rq_o[nxt][slot] = make_float4(no.x, no.y, no.z, kShadowEps);
            a.rq_d[nxt][slot] = Packed3{nd.x, nd.y, nd.z};
            a.rq_sid[nxt][slot] = sid;
        }
        if constexpr (GUIDED) wave_sync();   // the next vertex's network overwrites the lobes just read
    }
    if (GUIDED && __lane_id() == 0 && n_queries) atomicAdd(&a.stats[6], n_queries);
}

// The shading kernel for a render: standard surfaces / textures (STD), a
// guide (GUIDED), and the waves per SIMD (2, 3 or 4).
template <bool STD, bool GUIDED>
const void *shade_kernel_t(uint32_t waves) {
    if (waves >= 4) return reinterpret_cast<const void *>(&k_shade<4, STD, GUIDED>);
    if (waves == 3) return reinterpret_cast<const void *>(&k_shade<3, STD, GUIDED>);
    return reinterpret_cast<const void *>(&k_shade<2, STD, GUIDED>);
}
#endif  // FL_SHADE_TU

// -------------------------------------------------------------- resolve ----
// One thread per padded pixel of all items; sums its samples of this pass in
// sample order and adds them to the item's film.
__global__ void __launch_bounds__(kBlock) k_resolve(RenderArgs a, const float *cmf) {
    const uint32_t pg = blockIdx.x * kBlock + threadIdx.x;
    if (pg >= a.total_pix) return;
    const uint32_t item = a.n_items > 1 ? item_search(a, pg, [](const ItemDev &it) { return it.ppix_base; }) : 0u;
    const ItemDev &it = a.items[item];
    const uint32_t pl = pg - it.ppix_base;
    uint32_t tile = pl >> 6, within = pl & 63u;
    uint32_t x = (tile % it.tiles_x) * 8u + (within & 7u);
    uint32_t y = (tile / it.tiles_x) * 8u + (within >> 3);
    if (x >= it.cam.W || y >= it.cam.H) return;
    const int S = a.grid.stride;
    const float *cx = cmf, *cy = cmf + S, *cz = cmf + 2 * S;
    // samples are added to the film one at a time in sample order, starting
    // from the sums of the earlier passes: the film's bits then do not depend
    // on how the samples were split into passes (a batch item and the same
    // render alone split them differently)
    const size_t pix = (size_t)it.pix_base + y * it.cam.W + x;
    float sx = a.film_sum[3 * pix + 0], sy = a.film_sum[3 * pix + 1], sz = a.film_sum[3 * pix + 2];
    float qx = a.film_sq[3 * pix + 0], qy = a.film_sq[3 * pix + 1], qz = a.film_sq[3 * pix + 2];
    uint32_t nan_count = 0;
    for (uint32_t s = 0; s < it.spp_pass; ++s) {
        const uint32_t sid = it.path_base + s * it.npix + pl;
        float4 L4 = a.st_L[sid], p0 = a.pend[sid], p1 = a.pend[a.cap + sid], misc = a.st_misc[sid];
        float L[4] = {L4.x + p0.x + p1.x, L4.y + p0.y + p1.y, L4.z + p0.z + p1.z, L4.w + p0.w + p1.w};
        uint32_t flags = __float_as_uint(misc.w);
        Lanes ln = make_lanes(a.grid, misc.x);
        float X = 0.f, Y = 0.f, Z = 0.f;
        float w = a.xyz_weight;
        for (int k = 0; k < 4; ++k) {
            float wk = (flags & PF_SECTERM) ? (k == 0 ? 4.f * w : 0.f) : w;
            X += wk * spec_at(cx, ln, k) * L[k];
            Y += wk * spec_at(cy, ln, k) * L[k];
            Z += wk * spec_at(cz, ln, k) * L[k];
        }
        if (!isfinite(X) || !isfinite(Y) || !isfinite(Z)) { X = Y = Z = 0.f; ++nan_count; }
        sx += X; sy += Y; sz += Z;
        qx += X * X; qy += Y * Y; qz += Z * Z;
    }
    a.film_sum[3 * pix + 0] = sx;
    a.film_sum[3 * pix + 1] = sy;
    a.film_sum[3 * pix + 2] = sz;
    a.film_sq[3 * pix + 0] = qx;
    a.film_sq[3 * pix + 1] = qy;
    a.film_sq[3 * pix + 2] = qz;
    if (nan_count) atomicAdd(&a.stats[2], (unsigned long long)nan_count);
}

#ifdef FL_SHADE_TU
// ----------------------------------------------------------- BSDF queries --
// (compiled with the shading kernels, so that they run the same BSDF code)
__global__ void k_bsdf_query(const fl_bsdf_query *__restrict__ q, fl_bsdf_result *__restrict__ r, uint32_t n) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    fl_bsdf_query Q = q[i];
    Bsdf b;
    b.kind = Q.kind;
    b.ggx.ax = Q.alpha_x;
    b.ggx.ay = Q.alpha_y;
    b.weight = 1.f;
    b.vndf_pdf = (Q.flags & FL_RENDER_VNDF_PDF) != 0;
    for (int k = 0; k < 4; ++k) {
        b.s1[k] = Q.kind == MAT_CONDUCTOR ? Q.eta[k] : Q.r[k];
        b.s2[k] = Q.kind == MAT_CONDUCTOR ? Q.k[k] : Q.t[k];
    }
    b.ior = Q.eta[0];
    b.dispersive = Q.eta[1] != Q.eta[0] || Q.eta[2] != Q.eta[0] || Q.eta[3] != Q.eta[0];
    v3 wo(Q.wo[0], Q.wo[1], Q.wo[2]), wi(Q.wi[0], Q.wi[1], Q.wi[2]);
    fl_bsdf_result R;
    float f[4];
    bool refl = wi.z * wo.z > 0.f;  // ng = (0,0,1)
    bsdf_eval_nonspec(b, wo, wi, refl, f);
    for (int k = 0; k < 4; ++k) R.f_eval[k] = f[k];
    R.pdf_eval = bsdf_pdf_nonspec(b, wo, wi);
    BsdfSample s;
    if (wo.z == 0.f) {
        s.pdf = 0.f; s.type = 0; s.wi = v3(0.f, 0.f, 0.f);
        for (int k = 0; k < 4; ++k) s.f[k] = 0.f;
    } else {
        lobe_sample(b, wo, remap_u0(Q.u[0]), Q.u[1], s);
        if (s.pdf == 0.f) {
            s.type = 0;
            for (int k = 0; k < 4; ++k) s.f[k] = 0.f;
        } else if (!(s.type & BSDF_SPECULAR)) {
            bool rf = s.wi.z * wo.z > 0.f;
            if (!rf) for (int k = 0; k < 4; ++k) s.f[k] = 0.f;
            else if (b.kind == MAT_CONDUCTOR) microfacet_f(b, wo, s.wi, s.f, &s.wh);
        }
    }
    for (int k = 0; k < 4; ++k) R.f_sample[k] = s.f[k];
    R.pdf_sample = s.pdf;
    R.wi_sample[0] = s.wi.x; R.wi_sample[1] = s.wi.y; R.wi_sample[2] = s.wi.z;
    R.sampled_type = s.type;
    r[i] = R;
}

// Spectra of a standard-surface query: given per lane
struct QuerySpec {
    const fl_standard_query *q;
    __device__ float at(int which, int k) const {
        switch (which) {
        case SS_KD: return q->kd[k];
        case SS_KS: return q->ks[k];
        case SS_KT: return q->kt[k];
        case SS_SHEEN: return q->sheen[k];
        case SS_COAT: return q->coat[k];
        case SS_ETA: return q->eta[k];
        default: return q->k[k];
        }
    }
    __device__ float tf_lambda(int k) const { return q->lam[k]; }
};

// The StandardSurface BSDF for one material: evaluate and pdf with
// BSDF_ALL & ~BSDF_SPECULAR (what next-event estimation asks), sample with
// BSDF_ALL (what the path continuation asks), in the local frame (ng = z).
__global__ void k_bsdf_query_standard(StdParams m, const fl_standard_query *__restrict__ q, fl_bsdf_result *__restrict__ r,
                                      uint32_t n) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const fl_standard_query Q = q[i];
    const QuerySpec sp{&Q};
    const bool vndf = (Q.flags & FL_RENDER_VNDF_PDF) != 0;
    v3 wo(Q.wo[0], Q.wo[1], Q.wo[2]), wi(Q.wi[0], Q.wi[1], Q.wi[2]);
    fl_bsdf_result R;
    float f[4];
    std_bsdf_eval(m, sp, vndf, wo, wi, wi.z * wo.z > 0.f, kNonSpecular, f);
    for (int k = 0; k < 4; ++k) R.f_eval[k] = f[k];
    R.pdf_eval = std_bsdf_pdf(m, vndf, wo, wi, kNonSpecular);
    BsdfSample s;
    s.wi = v3(0.f, 0.f, 0.f);
    std_bsdf_sample(m, sp, vndf, wo, Q.u[0], Q.u[1], kAllLobes, s);
    if (s.pdf == 0.f) {
        s.type = 0;
        for (int k = 0; k < 4; ++k) s.f[k] = 0.f;
    }
    for (int k = 0; k < 4; ++k) R.f_sample[k] = s.f[k];
    R.pdf_sample = s.pdf;
    R.wi_sample[0] = s.wi.x; R.wi_sample[1] = s.wi.y; R.wi_sample[2] = s.wi.z;
    R.sampled_type = s.type;
    r[i] = R;
}

#endif  // FL_SHADE_TU

// The sampler of k_generate / k_shade for (pixel, sample, bounce) keys.
__global__ void k_sampler_draw(uint64_t seed, const uint32_t *__restrict__ keys, uint32_t n, uint32_t n_dims,
                               float *__restrict__ out) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (size_t)n * n_dims) return;
    const size_t k = i / n_dims;
    const uint32_t d = (uint32_t)(i - k * n_dims);
    out[i] = rng_u(rng_base(seed, keys[3 * k], keys[3 * k + 1], keys[3 * k + 2]), d);
}

template <typename T>
T *grow(void *&p, size_t &cap, size_t bytes, bool &err) {
    if (cap < bytes) {
        if (p) (void)hipFree(p);
        p = nullptr;
        cap = 0;
        if (hipMalloc(&p, bytes) != hipSuccess) { err = true; return nullptr; }
        cap = bytes;
    }
    return (T *)p;
}

} // namespace

namespace flx {
// The shading kernel for a render: standard surfaces / textures (STD), a
// guide (GUIDED), and the waves per SIMD (2, 3 or 4). Defined in the shading
// unit (this file compiled with FL_SHADE_TU).
const void *shade_kernel(bool std_, bool guided, uint32_t waves);
}

#ifdef FL_SHADE_TU
namespace flx {
const void *shade_kernel(bool std_, bool guided, uint32_t waves) {
    if (std_) return guided ? shade_kernel_t<true, true>(waves) : shade_kernel_t<true, false>(waves);
    return guided ? shade_kernel_t<false, true>(waves) : shade_kernel_t<false, false>(waves);
}

int bsdf_queries_standard(fl_ctx_t ctx, const fl_standard_surface *m, const fl_standard_query *d_q, fl_bsdf_result *d_r,
                          uint32_t n) {
    if (n == 0) return FL_OK;
    const StdParams P = standard_params(*m);
    hipLaunchKernelGGL(k_bsdf_query_standard, dim3((n + 255) / 256), dim3(256), 0, ctx->stream, P, d_q, d_r, n);
    FL_HIP_CHECK(ctx, hipGetLastError());
    return FL_OK;
}

int bsdf_queries(fl_ctx_t ctx, const fl_bsdf_query *d_q, fl_bsdf_result *d_r, uint32_t n) {
    if (n == 0) return FL_OK;
    hipLaunchKernelGGL(k_bsdf_query, dim3((n + 255) / 256), dim3(256), 0, ctx->stream, d_q, d_r, n);
    FL_HIP_CHECK(ctx, hipGetLastError());
    return FL_OK;
}
} // namespace flx
#else  // the main unit: the host driver

// ----------------------------------------------------------- host driver ---
struct RenderState {
    void *p[32] = {};
    size_t cap[32] = {};
    // pinned staging of the per-pass item table (two slots, each reused
    // only after the copy that read it has run)
    ItemDev *h_items[2] = {};
    uint32_t h_items_cap = 0;
    hipEvent_t item_ev[2] = {};
    uint32_t *h_counts = nullptr;     // pinned host copies of the queue counters, [2][16], then the record count
    hipEvent_t ev[2] = {};
    // bounce-loop graph (kChunk bounces), and the arguments it was captured with
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    RenderArgs gkey;
    int gvariant = -1;
    // films of the most recent render, resident in p[17] (sums): per item
    // its first pixel, size and samples per pixel
    struct Film { uint32_t pix_base, W, H, spp; };
    std::vector<Film> films;
};

namespace flx {

static RenderState *state_for(fl_ctx_t ctx) {
    if (!ctx->render_state) {
        RenderState *r = new (std::nothrow) RenderState();
        if (!r) return nullptr;
        memset(&r->gkey, 0, sizeof(r->gkey));
        if (hipHostMalloc((void **)&r->h_counts, 3 * 16 * sizeof(uint32_t), hipHostMallocDefault) != hipSuccess ||
            hipEventCreateWithFlags(&r->ev[0], hipEventDisableTiming) != hipSuccess ||
            hipEventCreateWithFlags(&r->ev[1], hipEventDisableTiming) != hipSuccess) {
            if (r->h_counts) (void)hipHostFree(r->h_counts);
            for (auto &e : r->ev)
                if (e) (void)hipEventDestroy(e);
            delete r;
            return nullptr;
        }
        ctx->render_state = r;
    }
    return ctx->render_state;
}

static void drop_graph(RenderState &b) {
    if (b.gexec) (void)hipGraphExecDestroy(b.gexec);
    if (b.graph) (void)hipGraphDestroy(b.graph);
    b.gexec = nullptr;
    b.graph = nullptr;
    b.gvariant = -1;
}

void release_render_buffers(fl_ctx_t ctx) {
    RenderState *b = ctx->render_state;
    if (!b) return;
    drop_graph(*b);
    for (int i = 0; i < 32; ++i)
        if (b->p[i]) (void)hipFree(b->p[i]);
    for (int k = 0; k < 2; ++k) {
        if (b->h_items[k]) (void)hipHostFree(b->h_items[k]);
        if (b->item_ev[k]) (void)hipEventDestroy(b->item_ev[k]);
    }
    if (b->h_counts) (void)hipHostFree(b->h_counts);
    for (auto &e : b->ev)
        if (e) (void)hipEventDestroy(e);
    delete b;
    ctx->render_state = nullptr;
}

// PhysicalCamera ctor with flipCameraDirection = true, pinhole (cameras/physical.h:21-43)
static CameraDev make_camera(const fl_camera *c) {
    auto dcross = [](const float a[3], const float b[3], float o[3]) {
        double ax = a[0], ay = a[1], az = a[2], bx = b[0], by = b[1], bz = b[2];
        o[0] = (float)(ay * bz - az * by);
        o[1] = (float)(az * bx - ax * bz);
        o[2] = (float)(ax * by - ay * bx);
    };
    auto norm = [](float v[3]) {
        float l = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
        v[0] /= l; v[1] /= l; v[2] /= l;
    };
    CameraDev cam;
    float pos[3] = {c->position[0], c->position[1], c->position[2]};
    float la[3] = {c->look_at[0], c->look_at[1], c->look_at[2]};
    float up[3] = {c->up[0], c->up[1], c->up[2]};
    float diff[3] = {pos[0] - la[0], pos[1] - la[1], pos[2] - la[2]};
    float fd = std::sqrt(diff[0] * diff[0] + diff[1] * diff[1] + diff[2] * diff[2]);
    float theta = c->vfov_deg * fl::kPi / 180.f;
    float hh = tanf(theta / 2.f);
    float aspect = (float)c->width / (float)c->height;
    float hw = aspect * hh;
    float w[3] = {diff[0], diff[1], diff[2]};
    norm(w);
    for (int k = 0; k < 3; ++k) w[k] = w[k] * -1.f;   // flipCameraDirection
    float u[3], v[3];
    dcross(up, w, u);
    norm(u);
    dcross(w, u, v);
    for (int k = 0; k < 3; ++k) w[k] = w[k] * -1.f;   // flip back
    for (int k = 0; k < 3; ++k) {
        cam.o[k] = pos[k];
        cam.llc[k] = pos[k] - (u[k] * hw * fd) - (v[k] * fd * hh) - (w[k] * fd);
        cam.hor[k] = u[k] * 2.f * hw * fd;
        cam.ver[k] = v[k] * 2.f * hh * fd;
    }
    cam.W = c->width;
    cam.H = c->height;
    return cam;
}

int render_items(fl_ctx_t ctx, const HostItem *hi, uint32_t n_items, const fl_render_params *p, fl_stats *stats) {
    hipStream_t st = ctx->stream;
    if (!p || n_items == 0 || !hi) return set_error(ctx, FL_E_INVALID, "fl_render: bad params");
    for (uint32_t k = 0; k < n_items; ++k) {
        if (!hi[k].scene || hi[k].scene->ctx != ctx)
            return set_error(ctx, FL_E_INVALID, "fl_render_batch: scene missing or of another context");
        if (hi[k].cam.width == 0 || hi[k].cam.height == 0 || hi[k].spp == 0)
            return set_error(ctx, FL_E_INVALID, "fl_render: bad camera or spp");
        if ((uint64_t)hi[k].cam.width * hi[k].cam.height > (1u << 26))
            return set_error(ctx, FL_E_INVALID, "fl_render: image larger than 2^26 pixels");
    }
    if (p->trav_prefetch > 3) return set_error(ctx, FL_E_INVALID, "fl_render: trav_prefetch must be 0, 1, 2 or 3");
    fl_guide_s *guide = (p->flags & FL_RENDER_GUIDED) ? p->guide : nullptr;
    if ((p->flags & FL_RENDER_GUIDED) && !guide) return set_error(ctx, FL_E_INVALID, "FL_RENDER_GUIDED needs params.guide");
    if (guide && guide->ctx != ctx) return set_error(ctx, FL_E_INVALID, "guide belongs to another context");
    for (uint32_t k = 1; guide && k < n_items; ++k)
        if (hi[k].scene != hi[0].scene)
            return set_error(ctx, FL_E_UNSUPPORTED, "guided batches render one scene (many cameras)");
    (void)hipGetLastError();  // do not report a stale error of an earlier call
    const int mode = p->spectral_mode == FL_SPECTRAL_PARITY ? 1 : 0;
    fl_scene_s *s = hi[0].scene;
    const SpecTables &tab = s->spec[mode];

    RenderArgs a;
    memset(&a, 0, sizeof(a));
    a.n_items = n_items;
    a.sc = TraceScene{s->d_nodes, s->d_tris};
    a.mats = s->d_mats;
    a.lights = s->d_lights;
    a.n_lights = s->n_lights;
    const bool nee_one = (p->flags & FL_RENDER_NEE_ONE) != 0;
    a.nee_all = s->n_lights <= 2 && !nee_one ? 1u : 0u;
    a.rgb_basis = s->rgb_basis;
    a.stds = s->d_std;
    a.tex = s->d_texscene;
    a.spec = tab.d_tab;
    a.grid = tab.grid;
    a.seed = p->seed;
    // items: cameras, padded pixel layers, film and padded-pixel bases
    std::vector<ItemDev> items(n_items);
    uint64_t total_pix = 0, film_pix = 0, samples = 0;
    uint32_t max_spp = 0, max_depth = 0;
    for (uint32_t k = 0; k < n_items; ++k) {
        ItemDev &it = items[k];
        memset(&it, 0, sizeof(it));
        fl_scene_s *sk = hi[k].scene;
        it.sc = TraceScene{sk->d_nodes, sk->d_tris};
        it.mats = sk->d_mats;
        it.lights = sk->d_lights;
        it.spec = sk->spec[mode].d_tab;
        it.n_lights = sk->n_lights;
        it.nee_all = sk->n_lights <= 2 && !nee_one ? 1u : 0u;
        it.rgb_basis = sk->rgb_basis;
        it.stds = sk->d_std;
        it.tex = sk->d_texscene;
        it.cam = make_camera(&hi[k].cam);
        it.tiles_x = (hi[k].cam.width + 7) / 8;
        it.npix = it.tiles_x * ((hi[k].cam.height + 7) / 8) * 64u;
        it.ppix_base = (uint32_t)total_pix;
        it.pix_base = (uint32_t)film_pix;
        it.seed = p->seed + (uint64_t)k * 0x9e3779b97f4a7c15ull;   // item 0 keeps the render's seed
        total_pix += it.npix;
        film_pix += (uint64_t)hi[k].cam.width * hi[k].cam.height;
        samples += (uint64_t)hi[k].cam.width * hi[k].cam.height * hi[k].spp;
        max_spp = std::max(max_spp, hi[k].spp);
        max_depth = std::max(max_depth, sk->depth);
        if (sk->spec[mode].grid.stride != tab.grid.stride)
            return set_error(ctx, FL_E_INVALID, "fl_render_batch: scenes with different spectral tables");
    }
    if (total_pix > (1u << 30)) return set_error(ctx, FL_E_INVALID, "fl_render_batch: too many pixels");
    a.total_pix = (uint32_t)total_pix;
    // paths in flight per pass: every persistent launch ends in a tail where
    // lanes run out of rays, so passes are made as large as memory allows
    // (248 B of queues and state per path; 2^26 paths = 16.6 GB), capped at
    // half the free device memory. Every item advances by the same number of
    // samples per pixel per pass.
    uint32_t max_paths = p->max_paths;
    if (!max_paths) {
        const size_t per_path = kBytesPerPath;
        size_t free_b = 0, total_b = 0;
        max_paths = 1u << 26;
        if (hipMemGetInfo(&free_b, &total_b) == hipSuccess) {
            const size_t fit = free_b / 2 / per_path;
            if (fit < max_paths) max_paths = (uint32_t)std::max<size_t>(fit, 1u << 20);
        }
    }
    if (max_paths < a.total_pix) max_paths = a.total_pix;
    uint32_t spp_pass = std::max(1u, std::min(max_spp, max_paths / a.total_pix));
    if ((uint64_t)a.total_pix * spp_pass * 2u >= (uint64_t)kVisClosest)   // shadow targets j * cap + id, tag bit 31
        return set_error(ctx, FL_E_INVALID, "fl_render: too many paths per pass (at most 2^30)");
    a.cap = a.total_pix * spp_pass;
    a.depth_d = p->diffuse_depth;
    a.depth_r = p->reflection_depth;
    a.depth_t = p->transmission_depth;
    a.rr_threshold = p->rr_threshold;
    a.rr_start = p->rr_start;
    a.flags = p->flags;
    a.xyz_weight = (tab.grid.hi - tab.grid.lo) / (4.f * 106.856895f);

    RenderState *Bp = state_for(ctx);
    if (!Bp) return set_error(ctx, FL_E_OOM, "render state");
    RenderState &B = *Bp;
    bool err = false;
    const size_t cap = a.cap;
    // traversal variant, then the persistent grid: as many resident blocks
    // as its occupancy (LDS stack + VGPRs) allows on every CU
    const int want_stack = p->stack_depth ? (int)p->stack_depth : kTraceVariants[0].stack;
    // trav_waves 0: the first variant listed with that stack depth
    const TraceVariant *tv = nullptr;
    for (const TraceVariant &v : kTraceVariants)
        if (!tv && v.stack == want_stack && (p->trav_waves == 0 || v.waves == (int)p->trav_waves)) tv = &v;
    if (!tv) return set_error(ctx, FL_E_INVALID, "fl_render: no traversal variant with this stack_depth/trav_waves");
    // Traversal mode, per pass (fl_render_params.trav_prefetch 0 = auto):
    // - scenes of short rays (all below kPrefetchTris triangles): one ray per
    //   thread up to 16 M paths (Cornell, 16 spp: 1.06 against 1.42 and
    //   1.79 ms of traversal), next rays loaded ahead beyond;
    // - scenes of long rays: next rays loaded ahead in passes from
    //   kPrefetchPaths paths; a smaller pass is mostly its launches' tails and
    //   cache refills, which the budgeted refill with static chunks shortens
    //   (a 1 M-path pass of the guided room: 10.3 ms, against 11.0 loading
    //   next rays ahead, 10.7 without static chunks and 13.4 one ray per
    //   thread).
    uint32_t max_tris = 0;
    for (uint32_t k = 0; k < n_items; ++k) max_tris = std::max(max_tris, hi[k].scene->n_tris);
    auto pass_mode = [&](uint32_t pass_paths) -> int {
        if (p->trav_prefetch == 1) return kTravPrefetch;
        if (p->trav_prefetch == 2) return kTravBudget;
        if (p->trav_prefetch == 3) return kTravFlat;
        if (max_tris < kPrefetchTris) return pass_paths <= kFlatPaths ? kTravFlat : kTravPrefetch;
        return pass_paths >= kPrefetchPaths ? kTravPrefetch : kTravBudget;
    };
    a.trav_mode = pass_mode(a.cap);
    // persistent grids of both instantiations: as many resident blocks as
    // their occupancy allows on every CU
    uint32_t grid_mode[2];
    for (int m = 0; m < 2; ++m) {
        int occ = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, tv->fn(m == kTravPrefetch, n_items > 1), kBlock, 0) != hipSuccess ||
            occ <= 0) {
            (void)hipGetLastError();
            occ = 4;
        }
        grid_mode[m] = (uint32_t)(ctx->n_cu * occ);
    }
    const uint32_t grid_pers = std::max(grid_mode[0], grid_mode[1]);   // sizes the spill area
    a.trav_budget = p->trav_budget ? (int)p->trav_budget : 16;
    a.refill_lanes = p->refill_lanes ? (int)std::min(64u, p->refill_lanes) : 0;   // 0: the mode's default
    size_t spill_per = (size_t)(max_depth > (uint32_t)tv->stack ? max_depth - tv->stack + 1 : 1);
    a.spill_stride = grid_pers * kBlock;
    // one-ray-per-thread passes index the spill area by ray (at most 3 per path)
    const bool flat_possible = p->trav_prefetch == 3 || (p->trav_prefetch == 0 && max_tris < kPrefetchTris);
    if (flat_possible) a.spill_stride = (uint32_t)std::max<size_t>(a.spill_stride, 3ull * cap);
    a.st_beta = grow<float4>(B.p[0], B.cap[0], cap * 16, err);
    a.st_L = grow<float4>(B.p[1], B.cap[1], cap * 16, err);
    a.st_misc = grow<float4>(B.p[2], B.cap[2], cap * 16, err);
    a.pend = grow<float4>(B.p[4], B.cap[4], 2 * cap * 16, err);
    a.rq_o[0] = grow<float4>(B.p[5], B.cap[5], cap * 16, err);
    a.rq_d[0] = grow<Packed3>(B.p[6], B.cap[6], cap * 12, err);
    a.rq_sid[0] = grow<uint32_t>(B.p[7], B.cap[7], cap * 4, err);
    a.rq_o[1] = grow<float4>(B.p[8], B.cap[8], cap * 16, err);
    a.rq_d[1] = grow<Packed3>(B.p[9], B.cap[9], cap * 12, err);
    a.rq_sid[1] = grow<uint32_t>(B.p[10], B.cap[10], cap * 4, err);
    // the hits' barycentrics are read only to look up textures
    bool has_tex = false;
    for (uint32_t k = 0; k < n_items; ++k) has_tex = has_tex || hi[k].scene->has_textures;
    a.hits = grow<float2>(B.p[11], B.cap[11], cap * 8, err);
    a.hit_uv = has_tex ? grow<float2>(B.p[26], B.cap[26], cap * 8, err) : nullptr;
    a.sq_o = grow<Packed3>(B.p[12], B.cap[12], 2 * cap * 12, err);
    a.sq_d = grow<Packed3>(B.p[13], B.cap[13], 2 * cap * 12, err);
    a.sq_tgt = grow<uint32_t>(B.p[14], B.cap[14], 2 * cap * 4, err);
    bool shadow_colors = false;
    for (uint32_t k = 0; k < n_items; ++k) shadow_colors = shadow_colors || hi[k].scene->shadow_colors;
    a.sq_occ = shadow_colors ? grow<float4>(B.p[25], B.cap[25], 2 * cap * 16, err) : nullptr;
    a.counts = grow<uint32_t>(B.p[15], B.cap[15], 64, err);
    a.stats = grow<unsigned long long>(B.p[16], B.cap[16], 256, err);
    a.film_sum = grow<float>(B.p[17], B.cap[17], (size_t)film_pix * 12, err);
    a.film_sq = grow<float>(B.p[18], B.cap[18], (size_t)film_pix * 12, err);
    a.items = grow<ItemDev>(B.p[23], B.cap[23], (size_t)n_items * sizeof(ItemDev), err);
    if (n_items > 1) a.st_item = grow<uint32_t>(B.p[24], B.cap[24], cap * 4, err);
    if (B.h_items_cap < n_items) {
        for (int k = 0; k < 2; ++k) {
            if (B.h_items[k]) (void)hipHostFree(B.h_items[k]);
            B.h_items[k] = nullptr;
            if (hipHostMalloc((void **)&B.h_items[k], (size_t)n_items * sizeof(ItemDev), hipHostMallocDefault) != hipSuccess)
                err = true;
        }
        B.h_items_cap = err ? 0u : n_items;
    }
    for (int k = 0; k < 2 && !err; ++k)
        if (!B.item_ev[k] && hipEventCreateWithFlags(&B.item_ev[k], hipEventDisableTiming) != hipSuccess) err = true;
    a.spill = grow<uint2>(B.p[19], B.cap[19], (size_t)a.spill_stride * spill_per * 8, err);
    if (err) return set_error(ctx, FL_E_OOM, "render buffers");

    // learned proposal: schedule, network pointers, lobes and training records
    const bool train = guide && (p->flags & FL_RENDER_TRAIN_GUIDE) && !guide->diverged;
    uint32_t train_spp = 0, train_rounds = 0;
    float *recs = nullptr;
    uint32_t *st_rec = nullptr, rec_cap = 0;
    if (guide) {
        a.guide_mix = p->guide_mix > 0.f && p->guide_mix < 1.f ? p->guide_mix : 0.5f;
        a.guide_bounces = p->guide_bounces ? p->guide_bounces : ~0u;
        a.gW = guide->d_w;
        a.gP = guide->d_params;
        for (int k = 0; k < 3; ++k) {
            const float pad = 0.01f * std::max(1e-6f, s->bmax[k] - s->bmin[k]);
            a.box_lo[k] = s->bmin[k] - pad;
            a.box_inv[k] = 1.f / (s->bmax[k] - s->bmin[k] + 2.f * pad);
        }
        if (train) {
            // (0.15 by default: on the attic room 0.25 reached plain's error in
            // 7% more time, the shorter training passes paying for the
            // network's slightly later start)
            const float fr = p->guide_train_fraction > 0.f && p->guide_train_fraction <= 1.f ? p->guide_train_fraction
                                                                                              : 0.15f;
            train_spp = std::min(max_spp, std::max(1u, (uint32_t)std::lround(fr * (float)max_spp)));
            train_rounds = std::min(guide->cfg.train_rounds, train_spp);
            const uint32_t round_spp = train_round_end(train_spp, train_rounds, train_rounds - 1) -
                                       train_round_end(train_spp, train_rounds, train_rounds - 2);
            rec_cap = (uint32_t)std::min<size_t>((size_t)a.total_pix * std::min(round_spp, spp_pass) * 6, 1u << 26);
            st_rec = grow<uint32_t>(B.p[21], B.cap[21], cap * 4, err);
            recs = grow<float>(B.p[22], B.cap[22], (size_t)rec_cap * 16 * sizeof(float), err);
        }
        if (err) return set_error(ctx, FL_E_OOM, "guide buffers");
    }

    const bool counters = (p->flags & FL_RENDER_COUNTERS) != 0;
    const bool timing = (p->flags & FL_RENDER_TIMING) != 0;
    hipEvent_t e0, e1, tev[4];
    FL_HIP_CHECK(ctx, hipEventCreate(&e0));
    FL_HIP_CHECK(ctx, hipEventCreate(&e1));
    for (int k = 0; k < 4; ++k) FL_HIP_CHECK(ctx, hipEventCreate(&tev[k]));
    auto blocks = [](size_t n) { return (unsigned)((n + kBlock - 1) / kBlock); };
    hipLaunchKernelGGL(k_zero_f32, dim3(blocks((size_t)film_pix * 3)), dim3(kBlock), 0, st, a.film_sum, (size_t)film_pix * 3);
    hipLaunchKernelGGL(k_zero_f32, dim3(blocks((size_t)film_pix * 3)), dim3(kBlock), 0, st, a.film_sq, (size_t)film_pix * 3);
    hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(64), 0, st, (uint32_t *)a.stats, 32u);
    FL_HIP_CHECK(ctx, hipEventRecord(e0, st));

    uint32_t done_spp = 0, passes = 0, max_iter_seen = 0, n_trace = 0, n_shade = 0;
    double ms_trace = 0.0, ms_shade = 0.0;
    const uint32_t max_iter = (uint32_t)std::max({p->diffuse_depth, p->reflection_depth, p->transmission_depth}) * 2 + 8;
    uint32_t host_counts[16];

    // Pipelined bounce loop: bounces are issued kChunk at a time with the
    // full persistent grids (every kernel reads its queue size on the
    // device); the host checks the queue sizes after chunk c while chunk c+1
    // already runs, so the GPU does not wait for a host round trip. With
    // FL_RENDER_GRAPH a chunk is one captured graph, replayed.
    constexpr uint32_t kChunk = 4;   // even: a chunk starts and ends on queue parity 0
    // standard surfaces and image textures run in the second shade kernel
    bool has_std = false;
    for (uint32_t k = 0; k < n_items; ++k)
        has_std = has_std || hi[k].scene->has_standard || hi[k].scene->has_textures;
    // shading waves per SIMD, 0 = automatic: 3 for the plain kernel (168
    // VGPRs, 2 spilled), 2 for the standard-surface kernel (346 spilled at 3)
    const uint32_t shade_waves = p->shade_waves ? p->shade_waves : (has_std ? 2u : 3u);
    const int variant = (counters ? 1 : 0) | (shade_waves >= 4 ? 2 : 0) | ((int)(tv - kTraceVariants) << 2) |
                        (has_std ? 1 << 16 : 0) | (shade_waves == 3 ? 1 << 17 : 0) | (guide ? 1 << 18 : 0);
    // Shading and inference run grid-stride loops in which every block takes
    // the same share of the queue, so their grids are exactly the blocks that
    // fit on the device at once: a grid of more blocks runs its remainder as
    // a second round at partial occupancy.
    auto resident = [&](const void *fn, int block, uint32_t fallback) {
        int occ = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, fn, block, 0) != hipSuccess || occ <= 0) {
            (void)hipGetLastError();
            return fallback;
        }
        return (uint32_t)ctx->n_cu * (uint32_t)occ;
    };
    const void *shade_fn = shade_kernel(has_std, guide != nullptr, shade_waves);
    const int shade_bs = shade_block(has_std, guide != nullptr);
    const uint32_t shade_grid_max = resident(shade_fn, shade_bs, (uint32_t)ctx->n_cu * (uint32_t)(8 * kBlock / shade_bs));
    // guided renders evaluate the network inside the shading kernel
    auto launch_shade = [&](const RenderArgs &args, dim3 g, int j) {
        RenderArgs ka = args;
        int kj = j;
        void *kargs[] = {&ka, &kj};
        (void)hipLaunchKernel(shade_fn, g, dim3(shade_bs), kargs, 0, st);
    };
    auto launch_bounce = [&](const RenderArgs &args, uint32_t j) {
        tv->launch(counters, dim3(grid_mode[args.trav_mode == kTravPrefetch ? 1 : 0]), st, args, (int)j);
        launch_shade(args, dim3(shade_grid_max), (int)j);
    };
    const bool use_graph = (p->flags & FL_RENDER_GRAPH) && !timing;
    // One captured chunk serves every pass with the same arguments: trace,
    // guide and shade read only scene, queue and state fields, so the
    // per-pass fields are cleared from the key.
    auto ensure_graph = [&]() -> int {
        RenderArgs key = a;
        key.n_paths = 0;
        key.n_cam_rays = 0;
        if (B.gexec && B.gvariant == variant && memcmp(&key, &B.gkey, sizeof(key)) == 0) return FL_OK;
        if (B.gexec) FL_HIP_CHECK(ctx, hipStreamSynchronize(st));   // no replay of the old graph still queued
        drop_graph(B);
        FL_HIP_CHECK(ctx, hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        for (uint32_t j = 0; j < kChunk; ++j) launch_bounce(key, j);
        hipGraph_t g = nullptr;
        hipError_t ce = hipStreamEndCapture(st, &g);
        if (ce != hipSuccess || !g) return set_error(ctx, FL_E_HIP, "bounce graph capture failed");
        B.graph = g;
        if (hipGraphInstantiate(&B.gexec, B.graph, nullptr, nullptr, 0) != hipSuccess) {
            drop_graph(B);
            return set_error(ctx, FL_E_HIP, "bounce graph instantiation failed");
        }
        B.gkey = key;
        B.gvariant = variant;
        return FL_OK;
    };

    double ms_guide = 0.0;
    uint32_t round = 0;
    uint32_t item_slot = 0;
    while (done_spp < max_spp) {
        uint32_t this_spp = std::min(spp_pass, max_spp - done_spp);
        const bool training_pass = train && done_spp < train_spp && !guide->diverged;
        if (training_pass) {
            // training passes of doubling size: each trains the guide that
            // samples the next, larger one
            const uint32_t end = train_round_end(train_spp, train_rounds, std::min(round, train_rounds - 1));
            this_spp = std::min(this_spp, std::max(1u, end > done_spp ? end - done_spp : 1u));
        }
        a.guided = (guide && !guide->diverged && !(training_pass && round == 0)) ? 1u : 0u;
        a.recs = training_pass ? recs : nullptr;
        a.st_rec = training_pass ? st_rec : nullptr;
        a.rec_count = training_pass ? a.counts + 8 : nullptr;
        a.rec_cap = training_pass ? rec_cap : 0u;
        // this pass's item table: every item advances by up to this_spp samples
        {
            ItemDev *hit = B.h_items[item_slot];
            FL_HIP_CHECK(ctx, hipEventSynchronize(B.item_ev[item_slot]));   // its last copy has run
            uint32_t path_base = 0, slot_base = 0;
            for (uint32_t k = 0; k < n_items; ++k) {
                ItemDev it = items[k];
                it.spp_pass = hi[k].spp > done_spp ? std::min(this_spp, hi[k].spp - done_spp) : 0u;
                it.sample_base = p->sample_offset + done_spp;
                it.path_base = path_base;
                it.slot_base = slot_base;
                path_base += it.npix * it.spp_pass;
                slot_base += it.cam.W * it.cam.H * it.spp_pass;
                hit[k] = it;
            }
            a.n_paths = path_base;
            a.n_cam_rays = slot_base;
            a.trav_mode = pass_mode(a.n_cam_rays);
            FL_HIP_CHECK(ctx, hipMemcpyAsync(const_cast<ItemDev *>(a.items), hit, (size_t)n_items * sizeof(ItemDev),
                                             hipMemcpyHostToDevice, st));
            FL_HIP_CHECK(ctx, hipEventRecord(B.item_ev[item_slot], st));
            item_slot ^= 1u;
        }
        if (use_graph) {
            const int grc = ensure_graph();
            if (grc != FL_OK) return grc;
        }
        hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(64), 0, st, a.counts, 16u);
        hipLaunchKernelGGL(k_generate, dim3(std::min<unsigned>(blocks(a.n_paths), ctx->n_cu * 16)), dim3(kBlock), 0, st, a);
        uint32_t ne = a.n_cam_rays, ns = 0;
        uint32_t iter = 0;
        if (!timing) {
            uint32_t issued = 0, checked = 0;
            const uint32_t max_chunks = (max_iter + kChunk - 1) / kChunk;
            for (;;) {
                while (issued < checked + 2 && issued < max_chunks) {
                    if (use_graph) FL_HIP_CHECK(ctx, hipGraphLaunch(B.gexec, st));
                    else
                        for (uint32_t j = 0; j < kChunk; ++j) launch_bounce(a, j);
                    FL_HIP_CHECK(ctx, hipMemcpyAsync(B.h_counts + 16 * (issued & 1), a.counts, 64,
                                                     hipMemcpyDeviceToHost, st));
                    FL_HIP_CHECK(ctx, hipEventRecord(B.ev[issued & 1], st));
                    ++issued;
                }
                if (checked == issued) break;
                FL_HIP_CHECK(ctx, hipEventSynchronize(B.ev[checked & 1]));
                const uint32_t *hc = B.h_counts + 16 * (checked & 1);
                ++checked;
                // after an even number of bounces the next trace reads these
                if (hc[C_EXT + 0] + hc[C_SH + 1] == 0) break;
            }
            // a chunk issued past the end finds empty queues; the stream
            // orders it before the next pass
            iter = checked * kChunk;
            n_trace += issued * kChunk;
            n_shade += issued * kChunk;
            ne = ns = 0;
        }
        for (; iter < max_iter && (ne + ns) > 0; ++iter) {
            const int cur = iter & 1;
            uint32_t tg = std::min<uint32_t>(grid_mode[a.trav_mode == kTravPrefetch ? 1 : 0], blocks((size_t)ne + ns));
            if (timing) FL_HIP_CHECK(ctx, hipEventRecord(tev[0], st));
            tv->launch(counters, dim3(tg), st, a, (int)iter);
            ++n_trace;
            if (timing) FL_HIP_CHECK(ctx, hipEventRecord(tev[1], st));
            if (ne > 0) {
                launch_shade(a, dim3(std::min<uint32_t>(shade_grid_max, (ne + shade_bs - 1) / shade_bs)), (int)iter);
                ++n_shade;
            }
            if (timing) FL_HIP_CHECK(ctx, hipEventRecord(tev[2], st));
            FL_HIP_CHECK(ctx, hipMemcpyAsync(host_counts, a.counts, 64, hipMemcpyDeviceToHost, st));
            FL_HIP_CHECK(ctx, hipStreamSynchronize(st));
            if (timing) {
                float m0 = 0.f, m1 = 0.f;
                (void)hipEventElapsedTime(&m0, tev[0], tev[1]);
                (void)hipEventElapsedTime(&m1, tev[1], tev[2]);
                ms_trace += m0;
                ms_shade += m1;
            }
            if (ne > 0) {
                ne = host_counts[C_EXT + (cur ^ 1)];
                ns = host_counts[C_SH + cur];
            } else {
                ne = 0;
                ns = 0;   // shadow-only trace: nothing follows
            }
        }
        max_iter_seen = std::max(max_iter_seen, iter);
        hipLaunchKernelGGL(k_resolve, dim3(blocks(a.total_pix)), dim3(kBlock), 0, st, a, a.spec);
        if (training_pass) {
            // finish this pass's records, then train on them before the next pass
            const auto t0 = std::chrono::steady_clock::now();
            uint32_t nrec = 0;
            FL_HIP_CHECK(ctx, hipMemcpyAsync(&B.h_counts[32], a.counts + 8, 4, hipMemcpyDeviceToHost, st));
            FL_HIP_CHECK(ctx, hipStreamSynchronize(st));
            nrec = std::min(B.h_counts[32], rec_cap);
            if (nrec) {
                hipLaunchKernelGGL(k_finish_records, dim3(blocks(nrec)), dim3(kBlock), 0, st, a, nrec);
                const int trc = guide_train_on_records(guide, recs, nrec, guide->cfg.steps_per_pass, guide->cfg.batch,
                                                       p->seed * 1315423911ull + round, st);
                if (trc != FL_OK) return trc;
            }
            ms_guide += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
            ++round;
        }
        done_spp += this_spp;
        ++passes;
    }
    FL_HIP_CHECK(ctx, hipEventRecord(e1, st));
    FL_HIP_CHECK(ctx, hipStreamSynchronize(st));
    FL_HIP_CHECK(ctx, hipGetLastError());
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, e0, e1);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    for (int k = 0; k < 4; ++k) (void)hipEventDestroy(tev[k]);
    B.films.resize(n_items);
    for (uint32_t k = 0; k < n_items; ++k)
        B.films[k] = RenderState::Film{items[k].pix_base, hi[k].cam.width, hi[k].cam.height, hi[k].spp};
    for (uint32_t k = 0; k < n_items; ++k) {
        fl_film *film = hi[k].film;
        if (!film || !film->xyz_sum) continue;
        const uint32_t W = hi[k].cam.width, H = hi[k].cam.height;
        if (film->width != W || film->height != H) return set_error(ctx, FL_E_INVALID, "film size mismatch");
        const size_t off = (size_t)items[k].pix_base * 3;
        FL_HIP_CHECK(ctx, hipMemcpy(film->xyz_sum, a.film_sum + off, (size_t)W * H * 12, hipMemcpyDeviceToHost));
        if (film->xyz_sqsum)
            FL_HIP_CHECK(ctx, hipMemcpy(film->xyz_sqsum, a.film_sq + off, (size_t)W * H * 12, hipMemcpyDeviceToHost));
    }
    if (stats) {
        unsigned long long h[12];
        FL_HIP_CHECK(ctx, hipMemcpy(h, a.stats, sizeof(h), hipMemcpyDeviceToHost));
        memset(stats, 0, sizeof(*stats));
        stats->ms_total = ms;
        stats->samples = samples;
        stats->rays_extension = h[0];
        stats->rays_shadow = h[1];
        stats->nan_samples = h[2];
        stats->nodes_visited = h[3];
        stats->tris_tested = h[4];
        stats->shadow_nodes_visited = h[7];
        stats->shadow_tris_tested = h[8];
        stats->stack_spills = h[5];
        stats->ms_extend = ms_trace;   // combined traversal kernel (extension + shadow rays)
        stats->ms_shade = ms_shade;
        stats->ms_guide = ms_guide;
        stats->ms_infer = 0.0;   // the network runs inside the shading kernel (ms_shade)
        stats->guide_queries = h[6];
        stats->launches_extend = n_trace;
        stats->launches_shade = n_shade;
        stats->passes = passes;
        stats->max_bounce = max_iter_seen;
        stats->paths_per_pass = a.cap;
        stats->stack_depth = (uint32_t)tv->stack;
    }
    return FL_OK;
}

int last_film(fl_ctx_t ctx, uint32_t item, const float **d_sum, uint32_t *W, uint32_t *H, uint32_t *spp) {
    RenderState *r = ctx->render_state;
    if (!r || item >= r->films.size() || !r->p[17])
        return set_error(ctx, FL_E_INVALID, "fl_develop_film: no such item in the last render");
    const RenderState::Film &f = r->films[item];
    *d_sum = static_cast<const float *>(r->p[17]) + (size_t)f.pix_base * 3;
    *W = f.W;
    *H = f.H;
    *spp = f.spp;
    return FL_OK;
}

int render(fl_scene_s *s, const fl_camera *camera, const fl_render_params *p, fl_film *film, fl_stats *stats) {
    if (!camera || !p) return set_error(s->ctx, FL_E_INVALID, "fl_render: null camera/params");
    HostItem h{s, *camera, film, p->spp};
    return render_items(s->ctx, &h, 1, p, stats);
}

// Coordinate descent over the render knobs with probe renders of the same
// items (films not downloaded, guide not trained). Each probe runs twice;
// the faster device time counts, so one-time costs (allocation, graph
// capture) do not decide.
int autotune_items(fl_ctx_t ctx, const HostItem *hi, uint32_t n, const fl_render_params *p0, fl_tune_result *out) {
    if (!p0 || !out || !hi || n == 0) return set_error(ctx, FL_E_INVALID, "fl_autotune: null argument");
    std::vector<HostItem> items(hi, hi + n);
    for (HostItem &h : items) {
        h.film = nullptr;
        h.spp = std::min(h.spp, 64u);
    }
    fl_render_params base = *p0;
    base.flags &= ~(uint32_t)(FL_RENDER_TIMING | FL_RENDER_COUNTERS | FL_RENDER_TRAIN_GUIDE);
    base.flags |= FL_RENDER_GRAPH;
    uint32_t tried = 0;
    auto probe = [&](const fl_render_params &q, double *ms) -> int {
        double best = 1e300;
        for (int r = 0; r < 2; ++r) {
            fl_stats st;
            const int rc = render_items(ctx, items.data(), n, &q, &st);
            if (rc != FL_OK) return rc;
            best = std::min(best, st.ms_total);
        }
        ++tried;
        *ms = best;
        return FL_OK;
    };
    double ms_default = 0.0;
    int rc = probe(base, &ms_default);
    if (rc != FL_OK) return rc;
    fl_render_params cur = base;
    double ms_cur = ms_default;
    // BVH node width: every scene of the job rebuilt at each width, then left
    // at the fastest
    std::vector<fl_scene_s *> scenes;
    for (const HostItem &h : items)
        if (std::find(scenes.begin(), scenes.end(), h.scene) == scenes.end()) scenes.push_back(h.scene);
    const uint32_t width0 = scenes[0]->bvh_width;
    uint32_t width_best = width0;
    auto rebuild_all = [&](uint32_t w) -> int {
        for (fl_scene_s *sc : scenes) {
            const int r = rebuild_bvh(ctx, sc, w);
            if (r != FL_OK) return r;
        }
        return FL_OK;
    };
    for (uint32_t w : {8u, 6u, 4u}) {
        bool same = true;
        for (fl_scene_s *sc : scenes) same = same && sc->bvh_width == w;
        if (same && w == width0) continue;   // the default probe measured it
        if ((rc = rebuild_all(w)) != FL_OK) return rc;
        double ms = 0.0;
        if ((rc = probe(cur, &ms)) != FL_OK) return rc;
        if (ms < ms_cur) { ms_cur = ms; width_best = w; }
    }
    if ((rc = rebuild_all(width_best)) != FL_OK) return rc;
    auto try_values = [&](auto set, const std::vector<uint32_t> &vals) -> int {
        fl_render_params keep = cur;
        for (uint32_t v : vals) {
            fl_render_params q = keep;
            set(q, v);
            double ms = 0.0;
            const int r = probe(q, &ms);
            if (r == FL_E_INVALID) { (void)hipGetLastError(); continue; }   // value not valid here
            if (r != FL_OK) return r;
            if (ms < ms_cur) { ms_cur = ms; cur = q; }
        }
        return FL_OK;
    };
    // paths in flight per pass (the largest fitting is the default)
    rc = try_values([](fl_render_params &q, uint32_t v) { q.max_paths = v; }, {1u << 22, 1u << 24});
    if (rc == FL_OK)   // traversal variant: (stack, waves) pairs of kTraceVariants
        for (const TraceVariant &v : kTraceVariants) {
            if (&v == &kTraceVariants[0]) continue;
            const uint32_t packed = (uint32_t)v.stack | ((uint32_t)v.waves << 16);
            rc = try_values([](fl_render_params &q, uint32_t pk) { q.stack_depth = pk & 0xffffu; q.trav_waves = pk >> 16; },
                            {packed});
            if (rc != FL_OK) break;
        }
    if (rc == FL_OK) rc = try_values([](fl_render_params &q, uint32_t v) { q.trav_budget = v; }, {8u, 32u});
    if (rc == FL_OK) rc = try_values([](fl_render_params &q, uint32_t v) { q.refill_lanes = v; }, {1u, 16u, 32u});
    if (rc == FL_OK) rc = try_values([](fl_render_params &q, uint32_t v) { q.trav_prefetch = v; }, {1u, 2u, 3u});
    if (rc == FL_OK)
        rc = try_values([](fl_render_params &q, uint32_t v) { q.flags = v ? (q.flags | FL_RENDER_SORT_SHADING) : q.flags; },
                        {1u});
    if (rc == FL_OK) rc = try_values([](fl_render_params &q, uint32_t v) { q.shade_waves = v; }, {2u, 3u, 4u});
    if (rc != FL_OK) return rc;
    memset(out, 0, sizeof(*out));
    out->max_paths = cur.max_paths;
    out->stack_depth = cur.stack_depth ? cur.stack_depth : (uint32_t)kTraceVariants[0].stack;
    out->trav_waves = cur.trav_waves;
    out->trav_budget = cur.trav_budget ? cur.trav_budget : 16u;
    out->refill_lanes = cur.refill_lanes;   // 0: the mode's default (24 loaded ahead, 16 budgeted)
    out->trav_prefetch = cur.trav_prefetch;
    out->shade_waves = cur.shade_waves;
    out->sort_shading = (cur.flags & FL_RENDER_SORT_SHADING) ? 1u : 0u;
    out->ms_best = ms_cur;
    out->ms_default = ms_default;
    out->n_tried = tried;
    out->node_width = width_best;
    return FL_OK;
}

int sampler_draw(fl_ctx_t ctx, uint64_t seed, const uint32_t *d_keys, uint32_t n, uint32_t n_dims, float *d_out) {
    const size_t total = (size_t)n * n_dims;
    hipLaunchKernelGGL(k_sampler_draw, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, ctx->stream, seed, d_keys, n,
                       n_dims, d_out);
    FL_HIP_CHECK(ctx, hipGetLastError());
    return FL_OK;
}

} // namespace flx
#endif  // FL_SHADE_TU
