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
// Memory access goes through a policy M (plain loads and stores on the host;
// agent-scope relaxed atomics on the device, where other threads wrote the
// subtree) so that the same code runs in a CPU check.
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

// Restructures the treelet rooted at internal node x. cost: 2^K floats;
// part: 2^K bytes (the caller's scratch), entry S at [S * ST] (a stride lets
// the threads of a block interleave their scratch in LDS).
template <int K, int ST, class M>
FL_HD void treelet_restructure(M &m, uint32_t x, float *cost, uint8_t *part) {
    uint32_t L[K];       // treelet leaves: internal node ids or primitive ids | leaf bit
    uint32_t I[K - 1];   // the treelet's internal nodes, I[0] = x
    int nl = 2, ni = 1;
    L[0] = m.left(x);
    L[1] = m.right(x);
    I[0] = x;
    float la[K];         // leaf areas
    for (int k = 0; k < 2; ++k) la[k] = m.area_of(L[k]);
    while (nl < K) {
        int best = -1;
        float best_a = -1.f;
        for (int k = 0; k < nl; ++k)
            if (!(L[k] & kTreeletLeafBit) && la[k] > best_a) {
                best_a = la[k];
                best = k;
            }
        if (best < 0) break;   // every leaf is a primitive
        const uint32_t c = L[best];
        I[ni++] = c;
        L[best] = m.left(c);
        la[best] = m.area_of(L[best]);
        L[nl] = m.right(c);
        la[nl] = m.area_of(L[nl]);
        ++nl;
    }
    // leaf boxes and subtree costs
    float lb[K][6];
    float lc[K];
    uint32_t lcount[K];
    for (int k = 0; k < nl; ++k) {
        m.box_of(L[k], lb[k]);
        lcount[k] = (L[k] & kTreeletLeafBit) ? 1u : m.count(L[k]);
        lc[k] = (L[k] & kTreeletLeafBit) ? 0.f : m.cost(L[k]);
    }
    const uint32_t full = (1u << nl) - 1u;
    // the optimal split of every subset of two or more leaves (subsets in
    // increasing order: each proper subset comes first)
    for (uint32_t S = 1; S <= full; ++S) {
        if ((S & (S - 1u)) == 0u) {   // one leaf
            int k = 0;
            while (!(S & (1u << k))) ++k;
            cost[S * ST] = lc[k];
            part[S * ST] = 0;
            continue;
        }
        float b[6] = {kInf, kInf, kInf, -kInf, -kInf, -kInf};
        uint32_t cnt = 0;
        for (int k = 0; k < nl; ++k)
            if (S & (1u << k)) {
                for (int e = 0; e < 3; ++e) {
                    b[e] = fminf(b[e], lb[k][e]);
                    b[3 + e] = fmaxf(b[3 + e], lb[k][3 + e]);
                }
                cnt += lcount[k];
            }
        const uint32_t low = S & (0u - S);
        float best = kInf;
        uint32_t bp = low;
        // proper non-empty subsets P of S holding S's lowest leaf (each split once)
        for (uint32_t P = (S - 1u) & S; P != 0u; P = (P - 1u) & S) {
            if (!(P & low)) continue;
            const float c = cost[P * ST] + cost[(S ^ P) * ST];
            if (c < best) {
                best = c;
                bp = P;
            }
        }
        cost[S * ST] = treelet_cost(b, cnt, best);
        part[S * ST] = (uint8_t)bp;
    }
    // rebuild top-down: subset S of the internal node `node`
    uint32_t stk_set[K], stk_node[K];
    int sp = 0, next = 1;
    stk_set[sp] = full;
    stk_node[sp] = x;
    ++sp;
    while (sp > 0) {
        --sp;
        const uint32_t S = stk_set[sp], node = stk_node[sp];
        const uint32_t P = part[S * ST], Q = S ^ P;
        uint32_t ch[2];
        const uint32_t sets[2] = {P, Q};
        for (int c = 0; c < 2; ++c) {
            const uint32_t T = sets[c];
            if ((T & (T - 1u)) == 0u) {   // one leaf: the treelet leaf itself
                int k = 0;
                while (!(T & (1u << k))) ++k;
                ch[c] = L[k];
            } else {
                ch[c] = I[next++];
                stk_set[sp] = T;
                stk_node[sp] = ch[c];
                ++sp;
            }
            if (ch[c] & kTreeletLeafBit) m.set_parent_leaf(ch[c] & ~kTreeletLeafBit, (int)node);
            else m.set_parent_int(ch[c], (int)node);
        }
        m.set_children(node, ch[0], ch[1]);
        float b[6] = {kInf, kInf, kInf, -kInf, -kInf, -kInf};
        uint32_t cnt = 0;
        for (int k = 0; k < nl; ++k)
            if (S & (1u << k)) {
                for (int e = 0; e < 3; ++e) {
                    b[e] = fminf(b[e], lb[k][e]);
                    b[3 + e] = fmaxf(b[3 + e], lb[k][3 + e]);
                }
                cnt += lcount[k];
            }
        m.set_node(node, b, cnt, cost[S * ST]);
    }
}

} // namespace fl
