// fl_trace.h -- per-lane traversal of the 8-wide quantized BVH.
//
// One ray per lane (wave64).  The traversal keeps a "node group" (a node's
// internal children still to visit, as a hit mask over octant-ordered bit
// positions) and a "triangle group" (hit leaf triangles of the last node), in
// the style of compressed wide BVH traversal (Ylitie, Karras, Laine, HPG
// 2017), with a short stack of node groups in LDS that spills to a per-lane
// global area when it is full.  Box tests take one fma per plane (the ray
// transformed into the node's grid once per node) and widen each slab
// interval by a bound on that arithmetic's rounding, so no box containing a
// hit is ever culled: the closest hit (with the smaller-ID tie rule) is then
// the same as a brute-force search.
//
// The traversal is a resumable state machine (TravState + trav_step with an
// iteration budget) so that persistent kernels can refill lanes whose ray has
// finished with new rays (one wave-aggregated atomic per refill) instead of
// idling until the slowest lane of the wave is done.
#pragma once
#include "fl_bvh.h"

namespace fl {

struct TraceScene {
    const Node8 *__restrict__ nodes;
    const TriRec *__restrict__ tris;
};

struct TraceHit {
    float t, u, v;
    uint32_t id;   // original triangle ID or kNoHit
    uint32_t rec;  // triangle record index (leaf order)
};

// pointer into LDS (address space 3), so stack accesses compile to
// ds_read/ds_write; an entry packs (group base, imask | hits << 8) in 64 bits
typedef __attribute__((address_space(3))) unsigned long long lds_u64;

struct LaneStack {
    lds_u64 *lds;      // lane's first entry; entry i at lds[i * lds_stride]
    int lds_stride;
    int cap;           // entries in LDS
    uint2 *spill;      // lane's first spill entry; entry i at spill[i * spill_stride]
    size_t spill_stride;
};

struct TravCounters {
    uint32_t nodes = 0, tris = 0, spills = 0;
    uint32_t any_nodes = 0, any_tris = 0;   // the share of any-hit rays (shadow rays to area lights)
};

__device__ __forceinline__ float safe_inv(float d) {
    const float eps = 1e-20f;
    float dd = fabsf(d) > eps ? d : copysignf(eps, d);
    return 1.0f / dd;
}

struct TravState {
    v3 o, d;
    float inv[3];
    float tmin, best_t, bu, bv;
    uint32_t best_id, best_rec;
    RayPre rp;
    uint32_t g_base, g_imask, g_hits, t_base, t_hits;
    int sp;
    // the direction's octant (bit k: negative along axis k), from the signs of
    // inv (never zero): recomputed where needed rather than held in a register
    __device__ __forceinline__ uint32_t oct() const {
        return (__float_as_uint(inv[0]) >> 31) | ((__float_as_uint(inv[1]) >> 30) & 2u) |
               ((__float_as_uint(inv[2]) >> 29) & 4u);
    }
};

__device__ __forceinline__ void trav_init(TravState &s, v3 o, v3 d, float tmin, float tmax) {
    s.o = o;
    s.d = d;
    s.inv[0] = safe_inv(d.x);
    s.inv[1] = safe_inv(d.y);
    s.inv[2] = safe_inv(d.z);
    s.rp = ray_pre(d);
    s.tmin = tmin;
    s.best_t = tmax;
    s.best_id = kNoHit;
    s.best_rec = 0;
    s.bu = s.bv = 0.f;
    s.g_base = 0;
    s.g_imask = 1u;
    s.g_hits = 1u << s.oct();  // the group holding the root (slot 0)
    s.t_base = 0;
    s.t_hits = 0;
    s.sp = 0;
}

// Runs at most `budget` loop iterations, and returns early (false) for all
// lanes once `leave` lanes of the wave (0: none) have finished, so that those
// start their next rays; returns true when the ray is done (state holds the
// closest hit, or with `any` the first accepted hit).
// `any` and `skip_lights` are per lane, so closest-hit and any-hit rays in one
// wave run the same instruction stream.
//
// Each iteration a lane does one costly step, a single triangle test or a
// single node test, and then, if it has nothing left to test, pops its next
// node group in the same iteration. A wave thus runs each of the two blocks
// once per iteration for all lanes that need it; draining a lane's whole
// triangle group in a loop instead ran that loop as often as the lane with
// the most triangles needed, with 3-4 of 64 lanes active on average, and a
// pop in an iteration of its own left the lane idle through the others'
// node tests (together 23% of the traversal time on the spheres workload).
template <bool COUNT>
__device__ __forceinline__ bool trav_step(const TraceScene &sc, TravState &s, const LaneStack &ls, int budget,
                                          TravCounters &cnt, bool any, bool skip_lights, int leave = 0) {
    // (signed: the lane counts are compared with a floor that may be negative)
    const int floor_active = leave > 0 ? (int)__popcll(__ballot(1)) - leave : -1;
    for (int it = 0; it < budget; ++it) {
        if (leave > 0 && it > 0 && (int)__popcll(__ballot(1)) <= floor_active) return false;   // enough lanes done
        if (s.t_hits) {
            uint32_t b = __ffs(s.t_hits) - 1;
            s.t_hits &= s.t_hits - 1u;
            uint32_t ri = s.t_base + b;
            const uint4 *tp = reinterpret_cast<const uint4 *>(sc.tris + ri);
            uint4 a = tp[0], bq = tp[1], c = tp[2];
            if (COUNT) {
                ++cnt.tris;
                if (any) ++cnt.any_tris;
            }
            if (!(skip_lights && c.w != 0u)) {
                v3 p0(__uint_as_float(a.x), __uint_as_float(a.y), __uint_as_float(a.z));
                v3 p1(__uint_as_float(bq.x), __uint_as_float(bq.y), __uint_as_float(bq.z));
                v3 p2(__uint_as_float(c.x), __uint_as_float(c.y), __uint_as_float(c.z));
                float t, u, v;
                if (tri_intersect(s.rp, s.o, p0, p1, p2, s.tmin, s.best_t, &t, &u, &v)) {
                    if (any) {
                        s.best_t = t; s.best_id = a.w; s.best_rec = ri; s.bu = u; s.bv = v;
                        return true;
                    }
                    if (closer(t, a.w, s.best_t, s.best_id)) {
                        s.best_t = t; s.best_id = a.w; s.best_rec = ri; s.bu = u; s.bv = v;
                    }
                }
            }
        } else if (s.g_hits) {
            uint32_t pos = __ffs(s.g_hits) - 1;
            s.g_hits &= s.g_hits - 1u;
            const uint32_t oct = s.oct();
            uint32_t slot = pos ^ oct;
            uint32_t nidx = s.g_base + __popc(s.g_imask & ((1u << slot) - 1u));
            if (s.g_hits) {
                uint2 e = make_uint2(s.g_base, s.g_imask | (s.g_hits << 8));
                if (s.sp < ls.cap) ls.lds[s.sp * ls.lds_stride] = ((unsigned long long)e.y << 32) | e.x;
                else { ls.spill[(size_t)(s.sp - ls.cap) * ls.spill_stride] = e; if (COUNT) ++cnt.spills; }
                ++s.sp;
            }
            const uint4 *np = reinterpret_cast<const uint4 *>(sc.nodes + nidx);
            uint4 w0 = np[0], w1 = np[1];
            // the ray's near plane per axis is lo (words 2..4) for a positive
            // direction and hi (words 5..7) for a negative one: each lane
            // loads the plane words in its own order (a per-lane address,
            // free since each lane reads its own node; swapping the twelve
            // words after the load took 24 selects per node, 2.7% of the
            // traversal time on the spheres scene)
            uint4 qn[3], qf[3];
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                const uint32_t neg = (oct >> k) & 1u;
                qn[k] = np[2 + k + 3 * neg];
                qf[k] = np[5 + k - 3 * neg];
            }
            if (COUNT) {
                ++cnt.nodes;
                if (any) ++cnt.any_nodes;
            }
            float org[3] = {__uint_as_float(w0.x), __uint_as_float(w0.y), __uint_as_float(w0.z)};
            float step[3] = {__uint_as_float((w0.w & 0xffu) << 23), __uint_as_float(((w0.w >> 8) & 0xffu) << 23),
                             __uint_as_float(((w0.w >> 16) & 0xffu) << 23)};
            uint32_t imask = w0.w >> 24;
            float oo[3] = {s.o.x, s.o.y, s.o.z};
            // Along axis k the ray meets grid plane q at
            //   t(q) = (org + q step - o) inv = q a + b,  a = step inv, b = (org - o) inv,
            // a exact (a power-of-two scale of inv). One fma per plane against
            // b -/+ e: the roundings of inv, b, b -/+ e and the fma together
            // stay below 2^-21 (65536 |a| + |b|) >= 2^-21 (|q a| + |b|), so
            // e = 2^-19 (65536 |a| + |b|) makes the near distance a lower and
            // the far distance an upper bound of the exact ones, and no box
            // holding a hit is culled. (e is per axis: a ray nearly parallel
            // to an axis has large |b| there, in that axis's units of t only.)
            float a3[3], bnear[3], bfar[3];
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                a3[k] = step[k] * s.inv[k];
                const float b = (org[k] - oo[k]) * s.inv[k];
                const float e = (fabsf(a3[k]) * 65536.0f + fabsf(b)) * 0x1p-19f;
                bnear[k] = b - e;
                bfar[k] = b + e;
            }
            uint32_t hits = 0;
#pragma unroll
            for (int q = 0; q < 8; ++q) {
                float tn3[3], tf3[3];
#pragma unroll
                for (int k = 0; k < 3; ++k) {
                    uint32_t wn = (q >> 1) == 0 ? qn[k].x : ((q >> 1) == 1 ? qn[k].y : ((q >> 1) == 2 ? qn[k].z : qn[k].w));
                    uint32_t wf = (q >> 1) == 0 ? qf[k].x : ((q >> 1) == 1 ? qf[k].y : ((q >> 1) == 2 ? qf[k].z : qf[k].w));
                    uint32_t qnear = (q & 1) ? (wn >> 16) : (wn & 0xffffu);
                    uint32_t qfar = (q & 1) ? (wf >> 16) : (wf & 0xffffu);
                    tn3[k] = __builtin_fmaf((float)qnear, a3[k], bnear[k]);
                    tf3[k] = __builtin_fmaf((float)qfar, a3[k], bfar[k]);
                }
                const float bn = fmaxf(fmaxf(tn3[0], tn3[1]), fmaxf(tn3[2], s.tmin));
                const float bf = fminf(fminf(tf3[0], tf3[1]), fminf(tf3[2], s.best_t));
                if (bn <= bf) hits |= 1u << q;
            }
            uint32_t inner = hits & imask;
            uint32_t leafs = hits & ~imask;
            // the internal children in visiting order, bit q moved to q ^ oct
            // by three conditional swaps (bits, pairs, nibbles): no loop over
            // the hit children, which ran as often as the lane with the most
            // hits needed (1.5% of the traversal time on the spheres scene)
            uint32_t ng = inner;
            ng = (oct & 1u) ? (((ng & 0x55u) << 1) | ((ng >> 1) & 0x55u)) : ng;
            ng = (oct & 2u) ? (((ng & 0x33u) << 2) | ((ng >> 2) & 0x33u)) : ng;
            ng = (oct & 4u) ? (((ng & 0x0fu) << 4) | ((ng >> 4) & 0x0fu)) : ng;
            uint32_t tm = 0;
            const uint32_t meta_lo = w1.z, meta_hi = w1.w;
            while (leafs) {
                uint32_t q = __ffs(leafs) - 1;
                leafs &= leafs - 1u;
                uint32_t m = (q < 4 ? (meta_lo >> (8 * q)) : (meta_hi >> (8 * (q - 4)))) & 0xffu;
                tm |= ((1u << (m >> 5)) - 1u) << (m & 31u);
            }
            s.g_base = w1.x;
            s.g_imask = imask;
            s.g_hits = ng;
            s.t_base = w1.y;
            s.t_hits = tm;
        }
        if (!s.t_hits && !s.g_hits) {   // next node group from the stack, or done
            if (s.sp == 0) return true;
            --s.sp;
            uint2 e;
            if (s.sp < ls.cap) {
                const unsigned long long v = ls.lds[s.sp * ls.lds_stride];
                e = make_uint2((uint32_t)v, (uint32_t)(v >> 32));
            } else {
                e = ls.spill[(size_t)(s.sp - ls.cap) * ls.spill_stride];
            }
            s.g_base = e.x;
            s.g_imask = e.y & 0xffu;
            s.g_hits = e.y >> 8;
        }
    }
    return false;
}

// Whole-ray traversal (no refill), for simple callers.
template <bool ANY, bool SKIP_LIGHTS, bool COUNT>
__device__ __forceinline__ bool traverse(const TraceScene &sc, v3 o, v3 d, float tmin, float tmax,
                                         const LaneStack &ls, TraceHit &hit, TravCounters &cnt) {
    TravState s;
    trav_init(s, o, d, tmin, tmax);
    while (!trav_step<COUNT>(sc, s, ls, 1 << 30, cnt, ANY, SKIP_LIGHTS)) {
    }
    hit.t = s.best_t;
    hit.u = s.bu;
    hit.v = s.bv;
    hit.id = s.best_id;
    hit.rec = s.best_rec;
    return s.best_id != kNoHit;
}

// Ray indices of a persistent traversal, handed out through a per-wave pool
// [pool_next, pool_end) from which the wave's lanes take consecutive indices
// as they need them (neighbouring queue entries are coherent rays). Wave w of
// the grid (W waves) first takes the 64 rays of chunk w, then, for a static
// share of the rays (static_share256/256), the chunks w + W, w + 2W, ... with
// no atomic; the rest, [dyn_base, n), comes from `fetch` (0 at launch)
// kChunk rays at a time. Every wave of the grid contends for that one
// counter, whose atomics serialise at the device's coherence point: at one
// atomic per 64 rays they took more time than the traversal's own memory
// traffic. Drawing from the shared counter keeps all waves on one advancing
// window of the queue, which large scenes need for their BVH to stay in the
// caches; static chunks let waves drift apart but cost no atomics, which
// serves small, cache-resident scenes.
// A draw takes kChunk rays while the queue holds many per wave, and fewer as
// it drains: the rays left at the counter (as of the wave's last draw) over
// twice the grid's waves, in multiples of 64 down to 64. With a fixed 256 a
// launch of few rays per wave (2 M rays over 5,120 waves) ended on a few waves
// that had each drawn 256 rays just before the counter ran out.
constexpr uint32_t kChunk = 256;
struct RayCursor {
    uint32_t n, nwaves, static_chunks, dyn_base;
    uint32_t next_static;           // the wave's next static chunk
    uint32_t pool_next, pool_end;   // wave-uniform
    uint32_t dyn_seen;              // the counter's value at the wave's last draw
    uint32_t *fetch;
    __device__ __forceinline__ RayCursor(uint32_t n_, uint32_t *fetch_, uint32_t static_share256)
        : n(n_), fetch(fetch_) {
        const uint32_t nthreads = gridDim.x * blockDim.x;
        nwaves = nthreads / 64u;
        // wave-uniform: held in scalar registers
        const uint32_t wave = __builtin_amdgcn_readfirstlane((blockIdx.x * blockDim.x + threadIdx.x) / 64u);
        const uint32_t rounds = (uint32_t)((uint64_t)n * static_share256 / 256u / nthreads);
        static_chunks = rounds * nwaves;
        dyn_base = rounds ? rounds * nthreads : nthreads;
        pool_next = wave * 64u;             // chunk w first, in every case
        pool_end = pool_next + 64u;
        next_static = wave + nwaves;
        dyn_seen = 0u;
    }
    // rays of the next dynamic draw (wave-uniform, a multiple of 64 in [64, kChunk])
    __device__ __forceinline__ uint32_t dyn_chunk() const {
        const uint32_t left = n > dyn_base + dyn_seen ? n - dyn_base - dyn_seen : 0u;
        const uint32_t c = (left / (2u * nwaves)) & ~63u;
        return c < 64u ? 64u : (c > kChunk ? kChunk : c);
    }
    // the next ray index of each lane with `need` (>= n: none left); call
    // with the whole wave
    __device__ __forceinline__ uint32_t take(bool need) {
        const unsigned long long m = __ballot(need);
        if (m == 0ull) return 0u;
        const uint32_t lane = __lane_id();
        const uint32_t want = (uint32_t)__popcll(m);
        const uint32_t avail = pool_end - pool_next;
        uint32_t nbase = 0, chunk = 0;
        if (avail < want) {
            if (next_static < static_chunks) {
                nbase = next_static * 64u;
                chunk = 64u;
                next_static += nwaves;
            } else {
                const int leader = __ffsll(m) - 1;
                chunk = dyn_chunk();
                if (lane == (uint32_t)leader) nbase = atomicAdd(fetch, chunk);
                dyn_seen = __builtin_amdgcn_readlane(nbase, leader) + chunk;
                nbase = dyn_base + dyn_seen - chunk;
            }
        }
        const uint32_t rank = (uint32_t)__popcll(m & ((1ull << lane) - 1ull));
        const uint32_t idx = rank < avail ? pool_next + rank : nbase + (rank - avail);
        if (avail < want) {
            pool_next = nbase + (want - avail);   // want - avail <= 64 <= chunk
            pool_end = nbase + chunk;
        } else {
            pool_next += want;
        }
        return idx;
    }
};

constexpr uint32_t kNoRay = ~0u;   // a lane without a ray

// Lane `src`'s value of v (every lane of the wave must take part).
__device__ __forceinline__ uint32_t lane_u32(uint32_t v, int src) {
    return (uint32_t)__builtin_amdgcn_ds_bpermute(src << 2, (int)v);
}
__device__ __forceinline__ float lane_f32(float v, int src) { return __uint_as_float(lane_u32(__float_as_uint(v), src)); }

// The end of a persistent traversal, once the wave's rays are all handed out:
// a free lane starts its own next ray, or takes over another lane's
// (prefetch mode: that lane is still busy with its current ray), and
// otherwise helps a lane still traversing. A lane with work to spare (two or
// more hit children in its node group, a stack entry, or one child and
// triangles) gives part of it to a free lane: the far half of the group's
// children, its top stack entry, or its one child. The helper loads the same
// ray from the queue and runs that part from the giver's current closest
// distance, with no hit of its own, and may give parts away in turn. A
// finished part goes back to the lane that gave it, which keeps the closer
// hit by the tie rule of `closer` (an any-hit ray: any hit). Every triangle
// lies in exactly one part and no part culls a box holding a hit at or below
// its bound, so the ray's result is the one a single lane finds. The owner of
// a ray stores it once its own part and all parts it gave away are done.
// Without this, a wave's last rays ran on a few lanes while the rest idled:
// on the spheres workload 5.7 of the 42 ms of traversal passed between the
// queue running out and the last wave's exit.
//
// Hand-overs move ray indices only (the receiver reloads the ray), so that
// little is live beside the traversal state. A finished part is handed back
// in the helper's own first three LDS stack entries (its stack is empty
// then, or abandoned after an any-hit).
template <bool COUNT, bool ANY_SKIP, class Src>
__device__ __forceinline__ void trace_drain(const Src &src, TraceScene sc, TravState &s, uint32_t ray, bool any,
                                            uint32_t nxt, const LaneStack &ls, TravCounters &cnt) {
    // traversal iterations between hand-outs (spheres: 2, 4 and 8 within 0.5%;
    // a 1-spp pass of the guided room: 10.3, 10.5 and 10.8 ms). In batches a
    // helper reloads its ray through the item table, and hand-outs every 8
    // iterations win (the 8-camera terrain at 64 spp: 688 Msamples/s, 639
    // at every 2, 664 without the drain).
    constexpr int kDrainBudget = Src::kMulti ? 8 : 2;
    const int lane = (int)__lane_id();
    const lds_u64 *wave_lds = ls.lds - lane;   // lane h's entry j: wave_lds[h + j * lds_stride]
    // `ray`: the queue index of the ray this lane works on (its own, or the one
    // it helps with)
    bool running = ray != kNoRay;   // the lane traverses a part of `ray`
    bool mine = running;            // the lane owns `ray` and stores it
    int parent = -1;                // helper: the lane that gave this part
    unsigned long long helpers = 0ull;   // lanes running parts this lane gave
    while (__ballot(running || mine || parent >= 0 || helpers != 0ull || nxt != kNoRay)) {
        // ---- loaded next rays: a free lane starts its own, or another lane's
        bool free = !running && !mine && parent < 0 && helpers == 0ull;
        unsigned long long I, P = __ballot(nxt != kNoRay);
        if (P) {
            uint32_t idx = kNoRay;
            if (free && nxt != kNoRay) {   // its own
                idx = nxt;
                nxt = kNoRay;
            }
            I = __ballot(free && idx == kNoRay);
            P = __ballot(nxt != kNoRay);
            int partner = -1;
            while (I && P) {   // wave-uniform
                const int h = __ffsll(I) - 1, g = __ffsll(P) - 1;
                I &= I - 1ull;
                P &= P - 1ull;
                if (lane == h || lane == g) partner = lane == h ? g : h;
            }
            const uint32_t other = lane_u32(nxt, partner >= 0 ? partner : lane);
            if (nxt != kNoRay && partner >= 0) nxt = kNoRay;   // given away
            if (free && idx == kNoRay && partner >= 0) idx = other;
            if (idx != kNoRay) {
                v3 o, d;
                float t0, t1;
                uint32_t anyw;
                src.load(idx, o, d, t0, t1, anyw, sc);
                ray = idx;
                any = (int32_t)anyw >= 0;
                trav_init(s, o, d, t0, t1);
                mine = running = true;
                free = false;
            }
        }
        // ---- parts of rays: the k-th free lane helps the k-th lane with work to spare
        const bool spare = running && (__popc(s.g_hits) >= 2 || s.sp > 0 || (s.g_hits != 0u && s.t_hits != 0u));
        I = __ballot(free);
        unsigned long long D = __ballot(spare);
        if (I && D) {
            int partner = -1;
            while (I && D) {   // wave-uniform
                const int h = __ffsll(I) - 1, g = __ffsll(D) - 1;
                I &= I - 1ull;
                D &= D - 1ull;
                if (lane == h || lane == g) partner = lane == h ? g : h;
            }
            uint32_t gb = 0u, gi = 0u, give = 0u;
            if (spare && partner >= 0) {   // the giver
                if (__popc(s.g_hits) >= 2) {
                    // keep the near half (the low bits), give the rest
                    uint32_t far = s.g_hits;
                    for (int k = (__popc(s.g_hits) + 1) / 2; k > 0; --k) far &= far - 1u;
                    gb = s.g_base;
                    gi = s.g_imask;
                    give = far;
                    s.g_hits ^= far;
                } else if (s.sp > 0) {
                    --s.sp;
                    uint2 e;
                    if (s.sp < ls.cap) {
                        const unsigned long long v = ls.lds[s.sp * ls.lds_stride];
                        e = make_uint2((uint32_t)v, (uint32_t)(v >> 32));
                    } else {
                        e = ls.spill[(size_t)(s.sp - ls.cap) * ls.spill_stride];
                    }
                    gb = e.x;
                    gi = e.y & 0xffu;
                    give = e.y >> 8;
                } else {
                    gb = s.g_base;
                    gi = s.g_imask;
                    give = s.g_hits;
                    s.g_hits = 0u;
                }
                helpers |= 1ull << partner;
            }
            const int from = (free && partner >= 0) ? partner : lane;
            const uint32_t hray = lane_u32(ray, from), hgb = lane_u32(gb, from), hgi = lane_u32(gi, from);
            const uint32_t hgive = lane_u32(give, from);
            const float tbest = lane_f32(s.best_t, from);
            if (free && partner >= 0) {
                v3 o, d;
                float t0, t1;
                uint32_t anyw;
                src.load(hray, o, d, t0, t1, anyw, sc);
                ray = hray;
                any = (int32_t)anyw >= 0;
                trav_init(s, o, d, t0, tbest);
                s.g_base = hgb;
                s.g_imask = hgi;
                s.g_hits = hgive;
                parent = partner;
                running = true;
            }
        }
        // ---- traverse
        if (running) {
            if (trav_step<COUNT>(sc, s, ls, kDrainBudget, cnt, any, any && ANY_SKIP)) running = false;
        }
        // ---- finished parts go back to the lanes that gave them
        const bool deliver = !running && helpers == 0ull && parent >= 0;
        if (deliver) {
            lds_u64 *r = ls.lds;
            r[0] = ((unsigned long long)s.best_id << 32) | __float_as_uint(s.best_t);
            r[ls.lds_stride] = ((unsigned long long)s.best_rec << 32) | __float_as_uint(s.bu);
            r[2 * ls.lds_stride] = __float_as_uint(s.bv);
        }
        const unsigned long long dep = __ballot(deliver);
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        unsigned long long got = helpers & dep;
        helpers &= ~dep;
        while (got) {
            const int h = __ffsll(got) - 1;
            got &= got - 1ull;
            const unsigned long long v0 = wave_lds[h];
            const uint32_t id = (uint32_t)(v0 >> 32);
            const float t = __uint_as_float((uint32_t)v0);
            if (id != kNoHit && (any ? s.best_id == kNoHit : closer(t, id, s.best_t, s.best_id))) {
                const unsigned long long v1 = wave_lds[h + ls.lds_stride], v2 = wave_lds[h + 2 * ls.lds_stride];
                s.best_t = t;
                s.best_id = id;
                s.best_rec = (uint32_t)(v1 >> 32);
                s.bu = __uint_as_float((uint32_t)v1);
                s.bv = __uint_as_float((uint32_t)v2);
            }
        }
        if (any && s.best_id != kNoHit) running = false;   // any-hit: one part's hit decides
        if (deliver) parent = -1;
        if (mine && !running && helpers == 0ull) {
            src.store(ray, s);
            mine = false;
        }
        // the next round's LDS writes come after every read of this round
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
}

// Persistent traversal over n rays with per-lane refill (the two modes are
// separate instantiations, PREFETCH, so each gets its own register budget).
// Src provides
//   void load(uint32_t i, v3 &o, v3 &d, float &tmin, float &tmax, uint32_t &anyw,
//             TraceScene &sc) const     (sc: the BVH this ray traverses; the ray
//             is any-hit when anyw's top bit is clear)
//   void store(uint32_t i, const TravState &s) const
//   TraceScene base_scene() const   (the scene of every ray, or of the first item)
// Rays come from a RayCursor (`fetch` must be 0 at launch). Two refill modes:
//  - prefetch: each lane holds its next ray already loaded, so a finished
//    lane starts it at once; the traversal yields to do that once
//    `refill_lanes` lanes of the wave have finished, and the next rays' loads
//    overlap the traversal that follows. Best for long rays (large scenes),
//    where the yields are rare against the traversal.
//  - refill at the end of `budget` iterations, once `refill_lanes` lanes are
//    idle (or all are), the refilled lanes waiting for their rays' loads.
//    Best for short rays, where the yields would cost more than idle lanes.
// All lanes of the wave must call this (it uses wave-wide ballots).
// `any` is per ray: the first hit ends it (emitters skipped when ANY_SKIP).
// scenes from this many triangles on traverse in prefetch mode by default
constexpr uint32_t kPrefetchTris = 65536;
// ... in passes of at least this many paths
constexpr uint32_t kPrefetchPaths = 4u << 20;
// scenes below kPrefetchTris trace one ray per thread in passes up to this many paths
constexpr uint32_t kFlatPaths = 16u << 20;

template <bool COUNT, bool ANY_SKIP, bool PREFETCH, class Src>
__device__ __forceinline__ void trace_persistent(const Src &src, uint32_t n, uint32_t *fetch, const LaneStack &ls,
                                                 int budget, int refill_lanes, TravCounters &cnt) {
    // a source whose rays all traverse one scene returns it from every load:
    // starting from it keeps the scene pointers uniform (scalar registers)
    TraceScene sc = src.base_scene();
    TravState s;
    RayCursor cur(n, fetch, PREFETCH ? 0u : 192u);   // static share: budgeted refill only
    const uint32_t first = cur.take(true);   // lane g of the grid: ray g
    uint32_t ray = kNoRay;
    bool any = false;
    bool exhausted = false;
    if (first < n) {
        ray = first;
        v3 o, d;
        float tmin, tmax;
        uint32_t anyw;
        src.load(first, o, d, tmin, tmax, anyw, sc);
        any = (int32_t)anyw >= 0;
        trav_init(s, o, d, tmin, tmax);
    } else {
        exhausted = true;
    }
    if constexpr (!PREFETCH) {
        // refill mode
        while (true) {
            if (__ballot(ray != kNoRay) == 0ull && __ballot(!exhausted) == 0ull) break;
            if (ray != kNoRay) {
                const bool done = trav_step<COUNT>(sc, s, ls, budget, cnt, any, any && ANY_SKIP);
                if (done) {
                    src.store(ray, s);
                    ray = kNoRay;
                }
            }
            // refill once at least refill_lanes lanes are idle (or all are): one
            // round of ray loads serves many lanes
            const bool need = (ray == kNoRay) && !exhausted;
            const unsigned long long m = __ballot(need);
            if (m && (__popcll(m) >= (unsigned)refill_lanes || __ballot(ray != kNoRay) == 0ull)) {
                const uint32_t idx = cur.take(need);
                if (need) {
                    if (idx >= n) {
                        exhausted = true;
                    } else {
                        ray = idx;
                        v3 o, d;
                        float tmin, tmax;
                        uint32_t anyw;
                        src.load(idx, o, d, tmin, tmax, anyw, sc);
                        any = (int32_t)anyw >= 0;
                        trav_init(s, o, d, tmin, tmax);
                    }
                }
            }
            if (__ballot(exhausted) != 0ull) {   // no ray left to hand out
                trace_drain<COUNT, ANY_SKIP>(src, sc, s, ray, any, kNoRay, ls, cnt);
                return;
            }
        }
        return;
    }
    // prefetch mode: the lane's next ray, loaded while the current one runs
    uint32_t nxt = kNoRay;
    v3 no(0.f, 0.f, 0.f), nd(0.f, 0.f, 0.f);
    float ntmin = 0.f, ntmax = 0.f;
    // the next ray's any-hit word as loaded (top bit clear: any-hit), tested
    // only when the ray starts so that no wait for the load comes earlier
    uint32_t nanyw = 0x80000000u;
    TraceScene nsc = sc;
    while (true) {
        const bool need = nxt == kNoRay && !exhausted;
        if (__ballot(need)) {
            const uint32_t idx = cur.take(need);
            if (need) {
                if (idx >= n) {
                    exhausted = true;
                } else {
                    nxt = idx;
                    src.load(idx, no, nd, ntmin, ntmax, nanyw, nsc);
                }
            }
        }
        if (ray == kNoRay && nxt != kNoRay) {
            ray = nxt;
            nxt = kNoRay;
            any = (int32_t)nanyw >= 0;
            sc = nsc;
            trav_init(s, no, nd, ntmin, ntmax);
        }
        if (__ballot(ray != kNoRay) == 0ull) break;   // no ray left to start
        if (ray != kNoRay) {
            const bool done = trav_step<COUNT>(sc, s, ls, budget, cnt, any, any && ANY_SKIP, refill_lanes);
            if (done) {
                src.store(ray, s);
                ray = kNoRay;
            }
        }
        if (__ballot(exhausted) != 0ull) {   // no ray left to hand out
            trace_drain<COUNT, ANY_SKIP>(src, sc, s, ray, any, nxt, ls, cnt);
            return;
        }
    }
}

} // namespace fl
