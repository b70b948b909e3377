// fl_treelet.h -- treelet restructuring of a binary BVH (Karras and Aila,
// "Fast Parallel Construction of High-Quality Bounding Volume Hierarchies",
// HPG 2013): at a node whose subtree is final, take the treelet of up to K
// leaves grown from it by repeatedly opening the leaf of largest surface
// area, find the binary topology over those leaves that minimises the summed
// surface area of its internal nodes (dynamic programming over the 2^K leaf
// subsets), and rebuild the treelet in place with the same internal node
// indices. Every internal node outside the treelet, the leaves and the
// primitive order stay as they were.
//
// The tree layout is the builder's: internal nodes 0..n-2, children in
// left/right (kLeafBit marks a primitive leaf, its index into the sorted
// primitive order), parents in parent_int / parent_leaf, boxes (lo.xyz,
// hi.xyz) in nbox, primitive counts in ncount, and ncost = the summed surface
// area of the subtree's internal nodes (the SAH cost up to constant factors
// and the constant primitive term).
//
// The work is split in three steps so that the lanes of a wave can share one
// treelet's dynamic program: treelet_form (one thread: grow the treelet),
// treelet_dp_size (the subsets of one size, spread over `nlanes` lanes; the
// sizes in increasing order, with a barrier between them) and treelet_rebuild
// (one thread: write the optimal topology); treelet_optimise runs the last
// two in one thread. Memory access goes through a policy M (plain loads and
// stores on the host; agent-scope relaxed atomics on the device, where other
// threads wrote the subtree), so that the same code runs in a CPU check.
#pragma once
#include <stdint.h>

#include "fl_math.h"

namespace fl {

constexpr uint32_t kTreeletLeafBit = 0x80000000u;
constexpr int kTreeletMax = 7;

FL_HD float box_half_area(const float b[6]) {
    const float dx = b[3] - b[0], dy = b[4] - b[1], dz = b[5] - b[2];
    return dx * dy + dy * dz + dz * dx;
}

// Cost of a subtree: the summed surface area of its internal nodes (the SAH
// cost up to constant factors; primitives cost nothing, their term being the
// same in every topology). A subtree with box b whose children cost `split`
// together costs area(b) + split. Models that price the 8-wide collapse's
// leaf slots (a subtree of at most 3 triangles tested as one leaf at the cost
// of a triangle test per triangle, an internal node at 0.5 to 4 of that) gave
// trees that were 1-2% slower to traverse on the spheres scene and 17% on
// the Cornell box.
FL_HD float treelet_cost(const float b[6], uint32_t c, float split) {
    (void)c;
    return box_half_area(b) + split;
}

// A treelet: its leaves (internal node ids, or primitive ids | leaf bit) with
// their boxes, subtree costs and primitive counts, and its internal nodes.
template <int K>
struct TreeletLeaves {
    uint32_t L[K];
    uint32_t I[K - 1];   // I[0] = the treelet's root
    float lb[K][6];
    float lc[K];
    uint32_t lcount[K];
    int nl;
};

// The dynamic program's table: per leaf subset S, the least cost of a binary
// tree over S and the subset P of S that forms its first child.
template <int K>
struct TreeletTable {
    float cost[1u << K];
    uint8_t part[1u << K];
    FL_HD float c(uint32_t S) const { return cost[S]; }
    FL_HD uint32_t p(uint32_t S) const { return part[S]; }
    FL_HD void set(uint32_t S, float cs, uint32_t ps) {
        cost[S] = cs;
        part[S] = (uint8_t)ps;
    }
};
// The same table with entry S at [S * ST] (the tables of a wave's lanes
// interleaved in LDS, so that lanes reading the same S hit different banks).
template <int ST>
struct TreeletTableStrided {
    float *cost;
    uint8_t *part;
    FL_HD float c(uint32_t S) const { return cost[S * ST]; }
    FL_HD uint32_t p(uint32_t S) const { return part[S * ST]; }
    FL_HD void set(uint32_t S, float cs, uint32_t ps) const {
        cost[S * ST] = cs;
        part[S * ST] = (uint8_t)ps;
    }
};

// a[i] and a[i] = v for a run-time i, as selects over compile-time indices:
// on the device the arrays then stay in registers (an indexed access would
// move the whole array to scratch memory).
template <int N, class T>
FL_HD T treelet_get(const T (&a)[N], int i) {
    T v = a[0];
#pragma unroll
    for (int k = 1; k < N; ++k)
        if (k == i) v = a[k];
    return v;
}
template <int N, class T>
FL_HD void treelet_put(T (&a)[N], int i, T v) {
#pragma unroll
    for (int k = 0; k < N; ++k)
        if (k == i) a[k] = v;
}

// Grows the treelet rooted at internal node x: starting from x's two
// children, open the internal leaf of largest surface area until there are K
// leaves or only primitives are left.
template <int K, class M>
FL_HD void treelet_form(M &m, uint32_t x, TreeletLeaves<K> &t) {
    uint32_t L[K];
    float la[K];
#pragma unroll
    for (int k = 0; k < K; ++k) {
        L[k] = kTreeletLeafBit;
        la[k] = -1.f;
    }
    L[0] = m.left(x);
    L[1] = m.right(x);
    t.I[0] = x;
    la[0] = m.area_of(L[0]);
    la[1] = m.area_of(L[1]);
    int nl = 2;
#pragma unroll
    for (int step = 1; step < K - 1; ++step) {   // each opening adds one leaf and one internal node
        int best = -1;
        float best_a = -1.f;
#pragma unroll
        for (int k = 0; k < K; ++k)
            if (k < nl && !(L[k] & kTreeletLeafBit) && la[k] > best_a) {
                best_a = la[k];
                best = k;
            }
        if (best < 0) break;   // every leaf is a primitive
        const uint32_t c = treelet_get(L, best);
        t.I[step] = c;
        const uint32_t a = m.left(c), b = m.right(c);
        treelet_put(L, best, a);
        treelet_put(la, best, m.area_of(a));
        treelet_put(L, nl, b);
        treelet_put(la, nl, m.area_of(b));
        ++nl;
    }
#pragma unroll
    for (int k = 0; k < K; ++k) {
        t.L[k] = L[k];
        if (k < nl) {
            m.box_of(L[k], t.lb[k]);
            const bool prim = (L[k] & kTreeletLeafBit) != 0;
            t.lcount[k] = prim ? 1u : m.count(L[k]);
            t.lc[k] = prim ? 0.f : m.cost(L[k]);
        }
    }
    t.nl = nl;
}

// The optimal split of every subset of sz leaves that this lane owns
// (subsets lane + 1, lane + 1 + nlanes, ...). Every subset of fewer leaves
// must be done (by all lanes) before.
template <int K, class T>
FL_HD void treelet_dp_size(const TreeletLeaves<K> &t, T &tab, int sz, uint32_t lane, uint32_t nlanes) {
    const int nl = t.nl;
    const uint32_t full = (1u << nl) - 1u;
    for (uint32_t S = lane + 1u; S <= full; S += nlanes) {
        if (__builtin_popcount(S) != sz) continue;
        if (sz == 1) {
            tab.set(S, treelet_get(t.lc, __builtin_ctz(S)), 0u);
            continue;
        }
        float b[6] = {kInf, kInf, kInf, -kInf, -kInf, -kInf};
        uint32_t cnt = 0;
#pragma unroll
        for (int k = 0; k < K; ++k)
            if (S & (1u << k)) {
                for (int e = 0; e < 3; ++e) {
                    b[e] = fminf(b[e], t.lb[k][e]);
                    b[3 + e] = fmaxf(b[3 + e], t.lb[k][3 + e]);
                }
                cnt += t.lcount[k];
            }
        // the proper subsets P of S that hold S's lowest leaf (each split
        // once), in decreasing order: P = low | Q, Q a proper subset of R
        const uint32_t low = S & (0u - S), R = S ^ low;
        float best = kInf;
        uint32_t bp = low;
        for (uint32_t Q = (R - 1u) & R;; Q = (Q - 1u) & R) {
            const uint32_t P = low | Q;
            const float c = tab.c(P) + tab.c(S ^ P);
            if (c < best) {
                best = c;
                bp = P;
            }
            if (Q == 0u) break;
        }
        tab.set(S, treelet_cost(b, cnt, best), bp);
    }
}

// Writes the optimal topology of the treelet rooted at x, top-down: subset S
// of internal node `node`, reusing the treelet's internal node indices.
template <int K, class T, class M>
FL_HD void treelet_rebuild(M &m, uint32_t x, const TreeletLeaves<K> &t, const T &tab) {
    const int nl = t.nl;
    uint32_t stk_set[K], stk_node[K];
    int sp = 0, next = 1;
    stk_set[sp] = (1u << nl) - 1u;
    stk_node[sp] = x;
    ++sp;
    while (sp > 0) {
        --sp;
        const uint32_t S = stk_set[sp], node = stk_node[sp];
        const uint32_t P = tab.p(S), Q = S ^ P;
        uint32_t ch[2];
        const uint32_t sets[2] = {P, Q};
        for (int c = 0; c < 2; ++c) {
            const uint32_t U = sets[c];
            if ((U & (U - 1u)) == 0u) {   // one leaf: the treelet leaf itself
                ch[c] = treelet_get(t.L, __builtin_ctz(U));
            } else {
                ch[c] = treelet_get(t.I, next++);
                stk_set[sp] = U;
                stk_node[sp] = ch[c];
                ++sp;
            }
            if (ch[c] & kTreeletLeafBit) m.set_parent_leaf(ch[c] & ~kTreeletLeafBit, (int)node);
            else m.set_parent_int(ch[c], (int)node);
        }
        m.set_children(node, ch[0], ch[1]);
        float b[6] = {kInf, kInf, kInf, -kInf, -kInf, -kInf};
        uint32_t cnt = 0;
#pragma unroll
        for (int k = 0; k < K; ++k)
            if (S & (1u << k)) {
                for (int e = 0; e < 3; ++e) {
                    b[e] = fminf(b[e], t.lb[k][e]);
                    b[3 + e] = fmaxf(b[3 + e], t.lb[k][3 + e]);
                }
                cnt += t.lcount[k];
            }
        m.set_node(node, b, cnt, tab.c(S));
    }
}

// The dynamic program and the rebuild of a formed treelet in one thread;
// `nlanes` only sets how the subsets are dealt out (the result does not
// depend on it).
template <int K, class T, class M>
FL_HD void treelet_optimise(M &m, uint32_t x, const TreeletLeaves<K> &t, T &tab, uint32_t nlanes = 1) {
    for (int sz = 1; sz <= t.nl; ++sz)
        for (uint32_t lane = 0; lane < nlanes; ++lane) treelet_dp_size<K>(t, tab, sz, lane, nlanes);
    treelet_rebuild<K>(m, x, t, tab);
}

// A treelet of at most kTreeletSmall leaves is a node of at most that many
// primitives (the formation opens every internal leaf below 7 leaves). Such a
// subtree ends up as the leaf slots of one 8-wide node, whatever its binary
// topology, so it keeps its topology; only its cost is recorded. (Optimising
// those too changed neither the nodes visited nor the traversal time on the
// spheres and room scenes.)
constexpr int kTreeletSmall = 4;
template <class M>
FL_HD void treelet_keep(M &m, uint32_t x) {
    float b[6];
    m.box_of(x, b);
    const uint32_t l = m.left(x), r = m.right(x);
    float c = box_half_area(b);
    if (!(l & kTreeletLeafBit)) c += m.cost(l);
    if (!(r & kTreeletLeafBit)) c += m.cost(r);
    m.set_cost(x, c);
}

} // namespace fl
