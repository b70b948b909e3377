// fl_build.hip -- GPU LBVH build collapsed to 8-wide, 16-bit quantized nodes.
//
// Replaces the Embree scene build (bvh/bvh.cpp:7-47: one RTCGeometry per
// triangle, rtcCommitScene).  Steps, all on the device:
//   1. triangle boxes + centroid bounds
//   2. 63-bit Morton codes of centroids, radix sort (hipCUB)
//   3. binary radix tree (Karras, HPG 2012), parent links
//   4. the internal nodes listed by depth (top-down, a launch per level),
//      then the bottom-up box refit, a launch per level
//   4b. two bottom-up sweeps of treelet restructuring, a launch per level
//      (fl_treelet.h: the SAH-optimal topology of every 7-leaf treelet,
//      Karras and Aila 2013; a wave's lanes share each treelet's dynamic
//      program)
//   5. top-down greedy collapse into 8-wide nodes: repeatedly open the child
//      with the largest surface area among those of more than 3 triangles
//      (among all, when none is that large); children with <= 3 triangles
//      become leaf slots; slots are ordered by octant so that the traversal visits
//      near children first (bit position = slot ^ ray octant)
//   6. 16-bit conservative quantization against a power-of-two grid whose
//      origin is a multiple of the step (planes dequantize exactly in fp32)
#include <hipcub/hipcub.hpp>
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>

#include "fl_internal.h"
#include "fl_treelet.h"

namespace {

constexpr uint32_t kLeafBit = 0x80000000u;
static_assert(kLeafBit == fl::kTreeletLeafBit, "one leaf bit for the builder and the treelet pass");
constexpr int kMaxLeafTris = 3;

__device__ __forceinline__ uint32_t ord(float f) {
    uint32_t u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float unord(uint32_t u) {
    return __uint_as_float((u & 0x80000000u) ? (u & 0x7fffffffu) : ~u);
}


// Triangle boxes (lo.xyz, hi.xyz) and centroid bounds: a grid-stride loop,
// the bounds reduced over the block and merged with one ordered-uint atomic
// per block and bound (an atomic per wave on the same six words serialised
// 41,000 waves: 2.8 ms for 2.6 M triangles).
constexpr int kBoxBlock = 256;
__global__ void __launch_bounds__(kBoxBlock) k_tri_boxes(const float *__restrict__ verts, const uint32_t *__restrict__ idx,
                                                         uint32_t n, float *__restrict__ box, uint32_t *cb) {
    float mn[3] = {fl::kInf, fl::kInf, fl::kInf}, mx[3] = {-fl::kInf, -fl::kInf, -fl::kInf};
    for (uint32_t i = blockIdx.x * kBoxBlock + threadIdx.x; i < n; i += gridDim.x * kBoxBlock) {
        const float *a = verts + 3ull * idx[3ull * i + 0];
        const float *b = verts + 3ull * idx[3ull * i + 1];
        const float *d = verts + 3ull * idx[3ull * i + 2];
        for (int k = 0; k < 3; ++k) {
            const float lo = fminf(a[k], fminf(b[k], d[k]));
            const float hi = fmaxf(a[k], fmaxf(b[k], d[k]));
            box[6ull * i + k] = lo;
            box[6ull * i + 3 + k] = hi;
            const float c = 0.5f * lo + 0.5f * hi;
            mn[k] = fminf(mn[k], c);
            mx[k] = fmaxf(mx[k], c);
        }
    }
    for (int off = 32; off >= 1; off >>= 1) {
        for (int k = 0; k < 3; ++k) {
            mn[k] = fminf(mn[k], __shfl_xor(mn[k], off));
            mx[k] = fmaxf(mx[k], __shfl_xor(mx[k], off));
        }
    }
    __shared__ float s_b[kBoxBlock / 64][6];
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0)
        for (int k = 0; k < 3; ++k) {
            s_b[w][k] = mn[k];
            s_b[w][3 + k] = mx[k];
        }
    __syncthreads();
    if (threadIdx.x < 6) {
        const int k = threadIdx.x;
        float v = s_b[0][k];
        for (int q = 1; q < kBoxBlock / 64; ++q) v = k < 3 ? fminf(v, s_b[q][k]) : fmaxf(v, s_b[q][k]);
        if (k < 3) atomicMin(&cb[k], ord(v));
        else atomicMax(&cb[k], ord(v));
    }
}

__device__ __forceinline__ uint64_t expand21(uint32_t v) {
    uint64_t x = v & 0x1fffffu;
    x = (x | (x << 32)) & 0x1f00000000ffffull;
    x = (x | (x << 16)) & 0x1f0000ff0000ffull;
    x = (x | (x << 8)) & 0x100f00f00f00f00full;
    x = (x | (x << 4)) & 0x10c30c30c30c30c3ull;
    x = (x | (x << 2)) & 0x1249249249249249ull;
    return x;
}

__global__ void k_morton(const float *__restrict__ box, uint32_t n, const uint32_t *cb,
                         uint64_t *keys, uint32_t *vals) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint32_t q[3];
    for (int k = 0; k < 3; ++k) {
        float lo = unord(cb[k]), hi = unord(cb[3 + k]);
        float c = 0.5f * box[6ull * i + k] + 0.5f * box[6ull * i + 3 + k];
        float ext = hi - lo;
        float x = ext > 0.f ? (c - lo) / ext : 0.f;
        x = fminf(fmaxf(x, 0.f), 1.f);
        q[k] = (uint32_t)(x * 2097151.0f);
    }
    keys[i] = (expand21(q[0]) << 2) | (expand21(q[1]) << 1) | expand21(q[2]);
    vals[i] = i;
}

__device__ __forceinline__ int delta(const uint64_t *keys, int n, int i, int j) {
    if (j < 0 || j >= n) return -1;
    uint64_t a = keys[i], b = keys[j];
    if (a == b) return 64 + __clz((uint32_t)(i ^ j));
    return __clzll(a ^ b);
}

// Binary radix tree over n sorted keys: internal nodes 0..n-2 (root 0).
__global__ void k_karras(const uint64_t *__restrict__ keys, int n, uint32_t *left, uint32_t *right,
                         int *parent_int, int *parent_leaf) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n - 1) return;
    int d = (delta(keys, n, i, i + 1) - delta(keys, n, i, i - 1)) > 0 ? 1 : -1;
    int dmin = delta(keys, n, i, i - d);
    int lmax = 2;
    while (delta(keys, n, i, i + lmax * d) > dmin) lmax <<= 1;
    int l = 0;
    for (int t = lmax >> 1; t >= 1; t >>= 1)
        if (delta(keys, n, i, i + (l + t) * d) > dmin) l += t;
    int j = i + l * d;
    int dnode = delta(keys, n, i, j);
    int s = 0, t = l;
    do {
        t = (t + 1) >> 1;
        if (delta(keys, n, i, i + (s + t) * d) > dnode) s += t;
    } while (t > 1);
    int gamma = i + s * d + (d < 0 ? d : 0);
    int lo = i < j ? i : j, hi = i < j ? j : i;
    uint32_t L = (lo == gamma) ? ((uint32_t)gamma | kLeafBit) : (uint32_t)gamma;
    uint32_t R = (hi == gamma + 1) ? ((uint32_t)(gamma + 1) | kLeafBit) : (uint32_t)(gamma + 1);
    left[i] = L;
    right[i] = R;
    if (L & kLeafBit) parent_leaf[L & ~kLeafBit] = i; else parent_int[L] = i;
    if (R & kLeafBit) parent_leaf[R & ~kLeafBit] = i; else parent_int[R] = i;
    if (i == 0) parent_int[0] = -1;
}

// The binary tree is processed level by level: a top-down breadth-first pass
// lists the internal nodes by depth (one launch per level), then the refit
// and each treelet sweep run one launch per level, deepest first. The nodes
// of one level head disjoint subtrees, and everything below them was written
// by earlier launches, so no launch needs atomics or fences between its
// threads. (Climbing from the leaves with a flag per node, as a refit commonly
// does, took an agent-scope acquire-release atomic per step; on this GPU,
// whose eight XCDs have L2 caches of their own, each of those writes back and
// invalidates the XCD's L2: 8.8 ms for the refit and 23 ms per treelet sweep
// on 2.6 M triangles.)

// Appends the internal children of the level's nodes to the next level's
// list (one atomic per wave).
__global__ void k_bfs_level(const uint32_t *__restrict__ level, uint32_t count, const uint32_t *__restrict__ left,
                            const uint32_t *__restrict__ right, uint32_t *next, uint32_t cap, uint32_t *cursor) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t ch[2] = {kLeafBit, kLeafBit};
    if (i < count) {
        const uint32_t x = level[i];
        ch[0] = left[x];
        ch[1] = right[x];
    }
    const bool in0 = !(ch[0] & kLeafBit), in1 = !(ch[1] & kLeafBit);
    const unsigned long long m0 = __ballot(in0), m1 = __ballot(in1);
    const uint32_t tot = (uint32_t)(__popcll(m0) + __popcll(m1));
    uint32_t base = 0;
    if (__lane_id() == 0 && tot) base = atomicAdd(cursor, tot);
    base = __shfl(base, 0);
    const unsigned long long lt = (1ull << __lane_id()) - 1ull;
    const uint32_t c0 = (uint32_t)__popcll(m0);
    const uint32_t s0 = base + (uint32_t)__popcll(m0 & lt), s1 = base + c0 + (uint32_t)__popcll(m1 & lt);
    // (a valid tree lists each internal node once; the bound only matters to
    // a broken one, which the host then reports)
    if (in0 && s0 < cap) next[s0] = ch[0];
    if (in1 && s1 < cap) next[s1] = ch[1];
}

// Refit of one level: box and primitive count from the two children.
__global__ void k_refit_level(const uint32_t *__restrict__ level, uint32_t count, const uint32_t *__restrict__ vals,
                              const float *__restrict__ tbox, const uint32_t *__restrict__ left,
                              const uint32_t *__restrict__ right, float *nbox, uint32_t *ncount) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    const uint32_t node = level[i];
    float lo[3] = {fl::kInf, fl::kInf, fl::kInf}, hi[3] = {-fl::kInf, -fl::kInf, -fl::kInf};
    uint32_t cnt = 0;
    const uint32_t ch[2] = {left[node], right[node]};
    for (int c = 0; c < 2; ++c) {
        const float *b = (ch[c] & kLeafBit) ? tbox + 6ull * vals[ch[c] & ~kLeafBit] : nbox + 6ull * ch[c];
        for (int k = 0; k < 3; ++k) {
            lo[k] = fminf(lo[k], b[k]);
            hi[k] = fmaxf(hi[k], b[3 + k]);
        }
        cnt += (ch[c] & kLeafBit) ? 1u : ncount[ch[c]];
    }
    for (int k = 0; k < 3; ++k) {
        nbox[6ull * node + k] = lo[k];
        nbox[6ull * node + 3 + k] = hi[k];
    }
    ncount[node] = cnt;
}

// The tree arrays for the treelet routine (fl_treelet.h): plain memory.
constexpr uint32_t kTreeletRounds = 2;
struct TreeletMem {
    uint32_t *l, *r;
    int *pint, *pleaf;
    float *nbox;
    uint32_t *ncount;
    float *ncost;
    const float *tbox;
    const uint32_t *vals;
    __device__ uint32_t left(uint32_t x) const { return l[x]; }
    __device__ uint32_t right(uint32_t x) const { return r[x]; }
    __device__ void box_of(uint32_t id, float b[6]) const {
        const float *t = (id & kLeafBit) ? tbox + 6ull * vals[id & ~kLeafBit] : nbox + 6ull * id;
        for (int e = 0; e < 6; ++e) b[e] = t[e];
    }
    __device__ float area_of(uint32_t id) const {
        float b[6];
        box_of(id, b);
        return fl::box_half_area(b);
    }
    __device__ float cost(uint32_t id) const { return ncost[id]; }
    __device__ uint32_t count(uint32_t id) const { return ncount[id]; }
    __device__ void set_parent_leaf(uint32_t p, int v) const { pleaf[p] = v; }
    __device__ void set_parent_int(uint32_t p, int v) const { pint[p] = v; }
    __device__ void set_children(uint32_t x, uint32_t a, uint32_t b) const {
        l[x] = a;
        r[x] = b;
    }
    __device__ void set_cost(uint32_t x, float cst) const { ncost[x] = cst; }
    __device__ void set_node(uint32_t x, const float b[6], uint32_t c, float cst) const {
        for (int e = 0; e < 6; ++e) nbox[6ull * x + e] = b[e];
        ncount[x] = c;
        ncost[x] = cst;
    }
};

// Treelet restructuring of one level, a wave per 64 nodes: every lane grows
// its node's treelet; treelets of at most kTreeletSmall leaves keep their
// topology, and the others are optimised one after another by the whole wave
// (each lane takes some of the leaf subsets of each size: 127 subsets for 7
// leaves), then written back by the lane that owns the treelet. (One thread
// per treelet running the whole dynamic program alone took about 2,000
// dependent LDS steps per treelet, with 40 KB of LDS per wave for every
// lane's table and so one wave per SIMD.)
constexpr int kTreeletBlock = 64;
__global__ void __launch_bounds__(kTreeletBlock) k_treelet_level(const uint32_t *__restrict__ level, uint32_t count,
                                                                 TreeletMem m) {
    constexpr int K = fl::kTreeletMax;
    __shared__ fl::TreeletLeaves<K> s_leaves;   // the treelet being optimised
    __shared__ fl::TreeletTable<K> s_tab;
    const uint32_t lane = threadIdx.x;
    const uint32_t i = blockIdx.x * kTreeletBlock + lane;
    const bool active = i < count;
    const uint32_t node = active ? level[i] : 0u;
    fl::TreeletLeaves<K> mine;
    if (active) fl::treelet_form<K>(m, node, mine);
    const bool small = active && mine.nl <= fl::kTreeletSmall;
    if (small) fl::treelet_keep(m, node);
    unsigned long long todo = __ballot(active && !small);
    while (todo) {
        const uint32_t j = (uint32_t)__ffsll(todo) - 1u;
        todo &= todo - 1ull;
        if (lane == j) s_leaves = mine;
        __syncthreads();
        const int nl = s_leaves.nl;
        for (int sz = 1; sz <= nl; ++sz) {
            fl::treelet_dp_size<K>(s_leaves, s_tab, sz, lane, kTreeletBlock);
            __syncthreads();
        }
        if (lane == j) fl::treelet_rebuild<K>(m, node, s_leaves, s_tab);
        __syncthreads();   // the next treelet overwrites the table
    }
}

// Levels of more than kTreeletWaveLevel nodes: every lane optimises its own
// node's treelet, with its table in LDS (the 64 lanes' tables interleaved:
// 40 KB per wave). Sharing each treelet among the wave's lanes pays only
// near the root, where a level has few nodes; on a full wave the lanes'
// treelets one after another took 20-40 us each.
constexpr uint32_t kTreeletWaveLevel = 8;
__global__ void __launch_bounds__(kTreeletBlock) k_treelet_level_lanes(const uint32_t *__restrict__ level,
                                                                       uint32_t count, TreeletMem m) {
    constexpr int K = fl::kTreeletMax;
    __shared__ float s_cost[kTreeletBlock << K];
    __shared__ uint8_t s_part[kTreeletBlock << K];
    const uint32_t lane = threadIdx.x;
    const uint32_t i = blockIdx.x * kTreeletBlock + lane;
    if (i >= count) return;
    const uint32_t node = level[i];
    fl::TreeletLeaves<K> mine;
    fl::treelet_form<K>(m, node, mine);
    if (mine.nl <= fl::kTreeletSmall) {
        fl::treelet_keep(m, node);
        return;
    }
    fl::TreeletTableStrided<kTreeletBlock> tab{s_cost + lane, s_part + lane};
    fl::treelet_optimise<K>(m, node, mine, tab);
}

struct BuildView {
    const float *verts;
    const uint32_t *idx, *mat, *light;
    const uint32_t *vals;
    const float *tbox;
    const uint32_t *left, *right;
    const float *nbox;
    const uint32_t *ncount;
    fl::Node8 *nodes;
    fl::TriRec *tris;
    uint32_t *counters;     // [0] nodes, [1] tris, [2] next work count
    int width;              // children per wide node, 2..8 (the node format holds 8)
};

__device__ void child_box(const BuildView &v, uint32_t c, float lo[3], float hi[3]) {
    const float *b;
    if (c & kLeafBit) b = v.tbox + 6ull * v.vals[c & ~kLeafBit];
    else b = v.nbox + 6ull * c;
    for (int k = 0; k < 3; ++k) { lo[k] = b[k]; hi[k] = b[3 + k]; }
}
__device__ uint32_t child_count(const BuildView &v, uint32_t c) {
    return (c & kLeafBit) ? 1u : v.ncount[c];
}

// exponent e with step 2^e: >= ext/65532, >= 2*ulp(maxabs), >= 2^-126
__device__ int quant_exponent(float lo, float hi) {
    double ext = (double)hi - (double)lo;
    int e = -126;
    if (ext > 0.0) {
        int E;
        double m = frexp(ext / 65532.0, &E); // x = m*2^E, m in [0.5,1)
        int ce = (m == 0.5) ? (E - 1) : E;  // ceil(log2 x)
        e = ce > e ? ce : e;
    }
    float maxabs = fmaxf(fabsf(lo), fabsf(hi));
    if (maxabs > 0.f) {
        int E;
        frexpf(maxabs, &E);                  // maxabs in [2^(E-1), 2^E)
        int eu = E - 23;                     // 2 * ulp(maxabs)
        e = eu > e ? eu : e;
    }
    return e > 127 ? 127 : e;
}

// One thread per work item: (binary node, wide node index).
__global__ void k_collapse(BuildView v, const uint32_t *__restrict__ work_in, uint32_t n_work,
                           uint32_t *work_out) {
    uint32_t w = blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= n_work) return;
    uint32_t bnode = work_in[2 * w];
    uint32_t widx = work_in[2 * w + 1];

    uint32_t ch[8];
    int cnt = 0;
    if (bnode & kLeafBit) {
        ch[cnt++] = bnode;
    } else {
        ch[cnt++] = v.left[bnode];
        ch[cnt++] = v.right[bnode];
        // Open the child of largest surface area, preferring children of more
        // than kMaxLeafTris triangles: a smaller one already fits one leaf
        // slot, and opening it would spend slots splitting a leaf instead of
        // flattening a larger subtree (0.5-1.5% faster traversal on the
        // spheres and room scenes, a third of the room's shadow-ray triangle
        // tests).
        while (cnt < v.width) {
            int best = -1;
            float bestA = -1.f;
            for (int pass = 0; pass < 2 && best < 0; ++pass)
            for (int c = 0; c < cnt; ++c) {
                if (ch[c] & kLeafBit) continue;
                if (pass == 0 && child_count(v, ch[c]) <= (uint32_t)kMaxLeafTris) continue;
                float lo[3], hi[3];
                child_box(v, ch[c], lo, hi);
                float dx = hi[0] - lo[0], dy = hi[1] - lo[1], dz = hi[2] - lo[2];
                float A = dx * dy + dy * dz + dz * dx;
                if (A > bestA) { bestA = A; best = c; }
            }
            if (best < 0) break;
            uint32_t b = ch[best];
            ch[best] = v.left[b];
            ch[cnt++] = v.right[b];
        }
    }

    // union box, centroids, slot assignment by octant
    float plo[3] = {fl::kInf, fl::kInf, fl::kInf}, phi[3] = {-fl::kInf, -fl::kInf, -fl::kInf};
    float clo[8][3], chi[8][3];
    for (int c = 0; c < cnt; ++c) {
        child_box(v, ch[c], clo[c], chi[c]);
        for (int k = 0; k < 3; ++k) {
            plo[k] = fminf(plo[k], clo[c][k]);
            phi[k] = fmaxf(phi[k], chi[c][k]);
        }
    }
    int slot_of[8];
    uint32_t used = 0;
    for (int c = 0; c < cnt; ++c) {
        int ideal = 0;
        for (int k = 0; k < 3; ++k) {
            float cc = 0.5f * clo[c][k] + 0.5f * chi[c][k];
            float pc = 0.5f * plo[k] + 0.5f * phi[k];
            if (cc > pc) ideal |= 1 << k;
        }
        int best = -1, bestd = 99;
        for (int s = 0; s < 8; ++s) {
            if (used & (1u << s)) continue;
            int dd = __popc((uint32_t)(s ^ ideal));
            if (dd < bestd) { bestd = dd; best = s; }
        }
        slot_of[c] = best;
        used |= 1u << best;
    }

    // classify, allocate
    uint32_t n_int = 0, n_tri = 0;
    uint32_t leafcnt[8];
    for (int c = 0; c < cnt; ++c) {
        uint32_t k = child_count(v, ch[c]);
        leafcnt[c] = k;
        if (k <= (uint32_t)kMaxLeafTris) n_tri += k; else n_int += 1;
    }
    uint32_t child_base = n_int ? atomicAdd(&v.counters[0], n_int) : 0u;
    uint32_t tri_base = n_tri ? atomicAdd(&v.counters[1], n_tri) : 0u;
    uint32_t work_base = n_int ? atomicAdd(&v.counters[2], n_int) : 0u;

    fl::Node8 node;
    int ex[3];
    for (int k = 0; k < 3; ++k) {
        ex[k] = quant_exponent(plo[k], phi[k]);
        double step = ldexp(1.0, ex[k]);
        double org = floor((double)plo[k] / step) * step;
        node.org[k] = (float)org;
    }
    node.ex = (uint8_t)(ex[0] + 127);
    node.ey = (uint8_t)(ex[1] + 127);
    node.ez = (uint8_t)(ex[2] + 127);
    node.imask = 0;
    node.child_base = child_base;
    node.tri_base = tri_base;
    for (int s = 0; s < 8; ++s) {
        node.meta[s] = 0;
        for (int k = 0; k < 3; ++k) { node.qlo[k][s] = 0xffff; node.qhi[k][s] = 0; }
    }
    // slot order for ranks / triangle offsets
    int child_in_slot[8];
    for (int s = 0; s < 8; ++s) child_in_slot[s] = -1;
    for (int c = 0; c < cnt; ++c) child_in_slot[slot_of[c]] = c;

    uint32_t rank = 0, toff = 0;
    for (int s = 0; s < 8; ++s) {
        int c = child_in_slot[s];
        if (c < 0) continue;
        for (int k = 0; k < 3; ++k) {
            double step = ldexp(1.0, ex[k]);
            double org = (double)node.org[k];
            double blo = (double)clo[c][k], bhi = (double)chi[c][k];
            if (leafcnt[c] > (uint32_t)kMaxLeafTris) {
                // an internal child's node quantizes its own slots on its own
                // grid (org_c, step_c, from the same box): they lie in
                // [org_c, org_c + min(65535, ceil((hi - org_c) / step_c) + 1) step_c].
                // Covering that range keeps every dequantized child box inside
                // this slot's box, not only the child's triangles.
                const int ec = quant_exponent(clo[c][k], chi[c][k]);
                const double sc = ldexp(1.0, ec);
                const double oc = floor(blo / sc) * sc;
                blo = oc;
                bhi = oc + fmin(ceil((bhi - oc) / sc) + 1.0, 65535.0) * sc;
            }
            double ql = floor((blo - org) / step) - 1.0;
            double qh = ceil((bhi - org) / step) + 1.0;
            ql = ql < 0.0 ? 0.0 : (ql > 65535.0 ? 65535.0 : ql);
            qh = qh < 0.0 ? 0.0 : (qh > 65535.0 ? 65535.0 : qh);
            node.qlo[k][s] = (uint16_t)ql;
            node.qhi[k][s] = (uint16_t)qh;
        }
        if (leafcnt[c] > (uint32_t)kMaxLeafTris) {
            node.imask |= (uint8_t)(1u << s);
            node.meta[s] = (uint8_t)rank;
            work_out[2 * (work_base + rank)] = ch[c];
            work_out[2 * (work_base + rank) + 1] = child_base + rank;
            ++rank;
        } else {
            node.meta[s] = (uint8_t)((toff & 31u) | (leafcnt[c] << 5));
            // gather the subtree's triangles (<= 3 leaves)
            uint32_t stk[8];
            int sp = 0;
            stk[sp++] = ch[c];
            while (sp) {
                uint32_t b = stk[--sp];
                if (b & kLeafBit) {
                    uint32_t t = v.vals[b & ~kLeafBit];
                    fl::TriRec r;
                    const uint32_t *ix = v.idx + 3ull * t;
                    for (int k = 0; k < 3; ++k) {
                        r.v0[k] = v.verts[3ull * ix[0] + k];
                        r.v1[k] = v.verts[3ull * ix[1] + k];
                        r.v2[k] = v.verts[3ull * ix[2] + k];
                    }
                    r.id = t;
                    r.material = v.mat ? v.mat[t] : 0u;
                    r.light = v.light[t];
                    v.tris[tri_base + toff] = r;
                    ++toff;
                } else {
                    stk[sp++] = v.right[b];
                    stk[sp++] = v.left[b];
                }
            }
        }
    }
    v.nodes[widx] = node;
}

__global__ void k_set_root_work(uint32_t *work, uint32_t root, uint32_t *counters) {
    work[0] = root;
    work[1] = 0;
    counters[0] = 1;  // root node allocated
    counters[1] = 0;
    counters[2] = 0;
}

__global__ void k_reset_u32(uint32_t *p) { *p = 0; }
// The root heads the level list; the next level is appended after it.
__global__ void k_level_init(uint32_t *order, uint32_t *cursor) {
    order[0] = 0u;
    *cursor = 1u;
}
constexpr size_t kMaxTreeLevels = 4096;

template <typename T>
int dalloc(fl_ctx_t ctx, T **p, size_t count) {
    hipError_t e = hipMalloc((void **)p, sizeof(T) * (count ? count : 1));
    if (e != hipSuccess)
        return flx::set_error(ctx, FL_E_OOM, std::string("hipMalloc: ") + hipGetErrorString(e));
    return FL_OK;
}

struct Scratch {
    std::vector<void *> ptrs;
    ~Scratch() {
        for (void *p : ptrs) (void)hipFree(p);
    }
};

} // namespace

namespace flx {

int build_bvh(fl_ctx_t ctx, fl_scene_s *s, const float *d_verts, const uint32_t *d_idx,
              const uint32_t *d_mat, const uint32_t *d_light) {
    const uint32_t n = s->n_tris;
    hipStream_t st = ctx->stream;
    (void)hipGetLastError();  // do not report a stale error of an earlier call
    if (n == 0) {   // empty scene: one root node without children, never hit
        fl::Node8 root;
        memset(&root, 0, sizeof(root));
        for (int k = 0; k < 3; ++k)
            for (int q = 0; q < 8; ++q) { root.qlo[k][q] = 0xffff; root.qhi[k][q] = 0; }
        root.ex = root.ey = root.ez = 127;
        if (hipMalloc((void **)&s->d_nodes, sizeof(fl::Node8)) != hipSuccess ||
            hipMalloc((void **)&s->d_tris, sizeof(fl::TriRec)) != hipSuccess)
            return set_error(ctx, FL_E_OOM, "hipMalloc empty scene");
        FL_HIP_CHECK(ctx, hipMemcpy(s->d_nodes, &root, sizeof(root), hipMemcpyHostToDevice));
        s->n_nodes = 1;
        s->depth = 1;
        s->build_ms = 0.0;
        return FL_OK;
    }
    Scratch sc;
#define ALLOC(p, cnt)                                                   \
    do {                                                                \
        int _r = dalloc(ctx, &p, (cnt));                                \
        if (_r) return _r;                                              \
        sc.ptrs.push_back((void *)p);                                   \
    } while (0)

    float *tbox; uint32_t *cb;
    uint64_t *keys_a, *keys_b; uint32_t *vals_a, *vals_b;
    uint32_t *left, *right; int *pint, *pleaf;
    float *nbox; uint32_t *ncount, *order, *work_a, *work_b, *counters;
    ALLOC(tbox, 6ull * n);
    ALLOC(cb, 6);
    ALLOC(keys_a, n); ALLOC(keys_b, n);
    ALLOC(vals_a, n); ALLOC(vals_b, n);
    uint32_t ni = n > 1 ? n - 1 : 1;
    ALLOC(left, ni); ALLOC(right, ni); ALLOC(pint, ni); ALLOC(pleaf, n);
    ALLOC(nbox, 6ull * ni); ALLOC(ncount, ni); ALLOC(order, ni);
    float *ncost;
    ALLOC(ncost, ni);
    ALLOC(work_a, 2ull * n + 2); ALLOC(work_b, 2ull * n + 2);
    ALLOC(counters, 4);

    hipEvent_t e0, e1;
    FL_HIP_CHECK(ctx, hipEventCreate(&e0));
    FL_HIP_CHECK(ctx, hipEventCreate(&e1));
    FL_HIP_CHECK(ctx, hipEventRecord(e0, st));

    const int B = 256;
    auto grid = [&](size_t m) { return (unsigned)((m + B - 1) / B); };
    uint32_t init_cb[6] = {0xffffffffu, 0xffffffffu, 0xffffffffu, 0u, 0u, 0u};
    FL_HIP_CHECK(ctx, hipMemcpyAsync(cb, init_cb, sizeof(init_cb), hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(k_tri_boxes, dim3(std::min(grid(n), 2048u)), dim3(kBoxBlock), 0, st, d_verts, d_idx, n, tbox, cb);
    hipLaunchKernelGGL(k_morton, dim3(grid(n)), dim3(B), 0, st, tbox, n, cb, keys_a, vals_a);

    size_t tmp_bytes = 0;
    FL_HIP_CHECK(ctx, hipcub::DeviceRadixSort::SortPairs(nullptr, tmp_bytes, keys_a, keys_b, vals_a,
                                                         vals_b, (int)n, 0, 63, st));
    void *tmp = nullptr;
    {
        uint8_t *t8;
        ALLOC(t8, tmp_bytes);
        tmp = t8;
    }
    FL_HIP_CHECK(ctx, hipcub::DeviceRadixSort::SortPairs(tmp, tmp_bytes, keys_a, keys_b, vals_a, vals_b,
                                                         (int)n, 0, 63, st));
    if (n >= 2) {
        hipLaunchKernelGGL(k_karras, dim3(grid(n - 1)), dim3(B), 0, st, keys_b, (int)n, left, right,
                           pint, pleaf);
        // internal nodes by depth: order[lstart[d] .. lstart[d + 1]) is level d
        std::vector<uint32_t> lstart;
        auto list_levels = [&]() -> int {
            hipLaunchKernelGGL(k_level_init, dim3(1), dim3(1), 0, st, order, counters + 3);
            lstart.assign({0u, 1u});
            for (;;) {
                const uint32_t b = lstart[lstart.size() - 2], e = lstart.back();
                hipLaunchKernelGGL(k_bfs_level, dim3(grid(e - b)), dim3(B), 0, st, order + b, e - b, left, right,
                                   order, n - 1, counters + 3);
                uint32_t cur = 0;
                FL_HIP_CHECK(ctx, hipMemcpyAsync(&cur, counters + 3, 4, hipMemcpyDeviceToHost, st));
                FL_HIP_CHECK(ctx, hipStreamSynchronize(st));
                if (cur == e) return FL_OK;
                if (cur < e || cur > n - 1 || lstart.size() > kMaxTreeLevels)
                    return set_error(ctx, FL_E_INVALID, "BVH build: binary tree levels inconsistent");
                lstart.push_back(cur);
            }
        };
        int rc = list_levels();
        if (rc) return rc;
        for (size_t d = lstart.size() - 1; d-- > 0;) {   // deepest level first
            const uint32_t b = lstart[d], c = lstart[d + 1] - lstart[d];
            hipLaunchKernelGGL(k_refit_level, dim3(grid(c)), dim3(B), 0, st, order + b, c, vals_b, tbox, left, right,
                               nbox, ncount);
        }
        // treelet restructuring of the binary tree (SAH over 7-leaf treelets).
        // A Morton order puts the few very large triangles of a scene (a ground
        // or light quad) deep in the tree, where they inflate every ancestor's
        // box: on the spheres scene the summed internal-node area falls 2.9x.
        const TreeletMem tm{left, right, pint, pleaf, nbox, ncount, ncost, tbox, vals_b};
        for (uint32_t r = 0; r < kTreeletRounds; ++r) {
            if (r > 0 && (rc = list_levels()) != FL_OK) return rc;   // the depths changed
            for (size_t d = lstart.size() - 1; d-- > 0;) {
                const uint32_t b = lstart[d], c = lstart[d + 1] - lstart[d];
                hipLaunchKernelGGL(c > kTreeletWaveLevel ? k_treelet_level_lanes : k_treelet_level,
                                   dim3((c + kTreeletBlock - 1) / kTreeletBlock), dim3(kTreeletBlock), 0, st, order + b,
                                   c, tm);
            }
        }
    }

    // device outputs (node capacity: one wide node per binary internal node + root)
    fl::Node8 *nodes;
    fl::TriRec *tris;
    {
        hipError_t e = hipMalloc((void **)&nodes, sizeof(fl::Node8) * (size_t)(ni + 1));
        if (e != hipSuccess) return set_error(ctx, FL_E_OOM, "hipMalloc nodes");
        e = hipMalloc((void **)&tris, sizeof(fl::TriRec) * (size_t)(n ? n : 1));
        if (e != hipSuccess) { (void)hipFree(nodes); return set_error(ctx, FL_E_OOM, "hipMalloc tris"); }
    }
    BuildView v;
    v.verts = d_verts; v.idx = d_idx; v.mat = d_mat; v.light = d_light;
    v.vals = vals_b; v.tbox = tbox; v.left = left; v.right = right;
    v.nbox = nbox; v.ncount = ncount; v.nodes = nodes; v.tris = tris; v.counters = counters;
    v.width = (int)s->bvh_width;

    uint32_t root = (n >= 2) ? 0u : (0u | kLeafBit);
    hipLaunchKernelGGL(k_set_root_work, dim3(1), dim3(1), 0, st, work_a, root, counters);
    uint32_t n_work = 1, depth = 0;
    uint32_t *win = work_a, *wout = work_b;
    while (n_work > 0) {
        if (++depth > 256) {
            (void)hipFree(nodes); (void)hipFree(tris);
            return set_error(ctx, FL_E_INVALID, "BVH collapse exceeded 256 levels");
        }
        hipLaunchKernelGGL(k_reset_u32, dim3(1), dim3(1), 0, st, counters + 2);
        hipLaunchKernelGGL(k_collapse, dim3(grid(n_work)), dim3(B), 0, st, v, win, n_work, wout);
        FL_HIP_CHECK(ctx, hipMemcpyAsync(&n_work, counters + 2, 4, hipMemcpyDeviceToHost, st));
        FL_HIP_CHECK(ctx, hipStreamSynchronize(st));
        uint32_t *t = win; win = wout; wout = t;
    }
    uint32_t cnts[2];
    FL_HIP_CHECK(ctx, hipMemcpyAsync(cnts, counters, 8, hipMemcpyDeviceToHost, st));
    FL_HIP_CHECK(ctx, hipEventRecord(e1, st));
    FL_HIP_CHECK(ctx, hipStreamSynchronize(st));
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, e0, e1);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);

    if (cnts[1] != n) {
        (void)hipFree(nodes); (void)hipFree(tris);
        return set_error(ctx, FL_E_INVALID, "BVH build lost triangles");
    }
    // shrink nodes to the exact count
    fl::Node8 *exact;
    if (hipMalloc((void **)&exact, sizeof(fl::Node8) * cnts[0]) == hipSuccess) {
        FL_HIP_CHECK(ctx, hipMemcpyAsync(exact, nodes, sizeof(fl::Node8) * cnts[0], hipMemcpyDeviceToDevice, st));
        FL_HIP_CHECK(ctx, hipStreamSynchronize(st));
        (void)hipFree(nodes);
        nodes = exact;
    }
    uint32_t h_cb[6];
    FL_HIP_CHECK(ctx, hipMemcpy(h_cb, cb, sizeof(h_cb), hipMemcpyDeviceToHost));
    s->d_nodes = nodes;
    s->d_tris = tris;
    s->n_nodes = cnts[0];
    s->depth = depth;
    s->build_ms = ms;
#undef ALLOC
    return FL_OK;
}

// Rebuilds the scene's BVH with `width` children per node from the build
// inputs the scene keeps on the device.
int rebuild_bvh(fl_ctx_t ctx, fl_scene_s *s, uint32_t width) {
    if (width < 2 || width > 8) return set_error(ctx, FL_E_INVALID, "BVH width must be 2..8");
    if (width == s->bvh_width) return FL_OK;
    FL_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));   // no render still reads the old nodes
    (void)hipFree(s->d_nodes);
    (void)hipFree(s->d_tris);
    s->d_nodes = nullptr;
    s->d_tris = nullptr;
    s->bvh_width = width;
    return build_bvh(ctx, s, s->d_verts, s->d_idx, s->d_mat, s->d_light);
}

} // namespace flx
