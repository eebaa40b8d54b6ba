// traverse.h -- closest-hit traversal of the 4-wide BVH (device code).
//
// One ray per lane.  The traversal stack keeps its top STACK_LDS entries in
// LDS (laid out [slot][lane] so the 64 lanes of a wave hit 64 different banks)
// and spills deeper entries to a per-lane private array.  Nodes are fetched as
// seven 16-byte loads from one 128-byte line; children are visited near-first.
// The loop is the while-while form with postponed leaves of Aila and Laine
// (2009): a lane that reaches its first leaf keeps it waiting and goes on
// through inner nodes until every lane of the wave has a leaf, so that node
// steps and triangle tests each run with most lanes of the wave busy.
#pragma once

#include "pppm_device.h"

#define STACK_LDS 16
#define STACK_OVF 112
#define STACK_MAX (STACK_LDS + STACK_OVF)

// Conservative slab test (Ize 2013): the exit distance is scaled up by
// 1 + 2^-22 so FP32 rounding cannot cull a box the triangle test would hit.
#define SLAB_FAR_SCALE 1.00000024f

// Per-lane traversal work, counted by the COUNT instantiation only.
struct TravCount {
    uint32_t nodes = 0;  // inner nodes visited (one 128-byte fetch each)
    uint32_t tris = 0;   // triangles tested (48 bytes each)
};

template <int BLOCK, bool COUNT = false>
__device__ inline Hit trace_closest(const SceneDev& sc, v3f o, v3f d, float tmin, float tmax,
                                    int32_t* __restrict__ stk, bool* overflow, TravCount* tc = nullptr) {
    Hit h;
    h.t = tmax;
    h.tri = -1;
    h.u = 0.0f;
    h.v = 0.0f;
    if (sc.root == EMPTY_REF) return h;
    const RayPre pre = ray_pre(o, d);
    // inverse direction as the reference computes it (qbvh_accel.cc:219-221)
    const float ix = d.x == 0.0f ? 1.0e20f : 1.0f / d.x;
    const float iy = d.y == 0.0f ? 1.0e20f : 1.0f / d.y;
    const float iz = d.z == 0.0f ? 1.0e20f : 1.0f / d.z;
    const float ox = o.x * ix, oy = o.y * iy, oz = o.z * iz;
    int32_t ovf[STACK_OVF];
    int sp = 0;
    // cur: an inner node (>= 0), a leaf (< 0) or EMPTY_REF (the stack is
    // empty); leaf: a leaf waiting for its triangle tests (< 0) or none
#define POP() (sp == 0 ? (int32_t)EMPTY_REF : (sp--, sp < STACK_LDS ? stk[sp * BLOCK] : ovf[sp - STACK_LDS]))
    int32_t cur = sc.root, leaf = EMPTY_REF;
    if (cur < 0) {
        leaf = cur;
        cur = EMPTY_REF;
    }
    while (true) {
        // inner nodes, until every lane of the wave has a leaf waiting
        while ((uint32_t)cur < (uint32_t)EMPTY_REF) {
            if (COUNT) tc->nodes++;
            const float4* np = reinterpret_cast<const float4*>(sc.nodes + cur);
            const float4 lx = np[0], hx = np[1], ly = np[2], hy = np[3], lz = np[4], hz = np[5];
            const int4 ch = reinterpret_cast<const int4*>(np)[6];
            // one 32-bit key per child: the entry distance (non-negative, so
            // its bit pattern orders like the value) with the child slot in the
            // two low mantissa bits; a missed child keys above every distance
            uint32_t key[4];
            const float lxa[4] = {lx.x, lx.y, lx.z, lx.w}, hxa[4] = {hx.x, hx.y, hx.z, hx.w};
            const float lya[4] = {ly.x, ly.y, ly.z, ly.w}, hya[4] = {hy.x, hy.y, hy.z, hy.w};
            const float lza[4] = {lz.x, lz.y, lz.z, lz.w}, hza[4] = {hz.x, hz.y, hz.z, hz.w};
#pragma unroll
            for (int c = 0; c < 4; c++) {
                float x0 = lxa[c] * ix - ox, x1 = hxa[c] * ix - ox;
                float y0 = lya[c] * iy - oy, y1 = hya[c] * iy - oy;
                float z0 = lza[c] * iz - oz, z1 = hza[c] * iz - oz;
                float n = fmaxf(fmaxf(fminf(x0, x1), fminf(y0, y1)), fmaxf(fminf(z0, z1), tmin));
                float f = fminf(fminf(fmaxf(x0, x1), fmaxf(y0, y1)), fminf(fmaxf(z0, z1), h.t));
                key[c] = (n <= f * SLAB_FAR_SCALE) ? ((__float_as_uint(n) & 0x7FFFFFFCu) | (uint32_t)c)
                                                   : (0xFFFFFFFCu | (uint32_t)c);
            }
            // sorting network, ascending, on min / max only (near child first)
            uint32_t k0 = key[0], k1 = key[1], k2 = key[2], k3 = key[3];
    #define KSORT(a, b)                  \
        {                                \
            const uint32_t lo = min(a, b); \
            b = max(a, b);               \
            a = lo;                      \
        }
            KSORT(k0, k1);
            KSORT(k2, k3);
            KSORT(k0, k2);
            KSORT(k1, k3);
            KSORT(k1, k2);
    #undef KSORT
            const int32_t ref[4] = {ch.x, ch.y, ch.z, ch.w};
            const uint32_t MISS = 0x7F800000u;
            int32_t push[3] = {ref[k3 & 3u], ref[k2 & 3u], ref[k1 & 3u]};
            bool pv[3] = {k3 < MISS, k2 < MISS, k1 < MISS};
#pragma unroll
            for (int q = 0; q < 3; q++) {
                if (pv[q]) {
                    if (sp < STACK_LDS) {
                        stk[sp * BLOCK] = push[q];
                    } else if (sp < STACK_MAX) {
                        ovf[sp - STACK_LDS] = push[q];
                    } else {
                        *overflow = true;
                        sp--;
                    }
                    sp++;
                }
            }
            cur = k0 < MISS ? ref[k0 & 3u] : POP();
            if (cur < 0 && leaf >= 0) {  // the first leaf waits; traversal goes on
                leaf = cur;
                cur = POP();
            }
            if (__ballot(leaf >= 0) == 0) break;
        }
        // the waiting leaf, then a leaf popped behind it
        while (leaf < 0) {
            const uint32_t v = (uint32_t)(~leaf);
            const int32_t first = (int32_t)(v >> 3);
            const int32_t cnt = (int32_t)(v & 7u) + 1;
            if (COUNT) tc->tris += (uint32_t)cnt;
            for (int k = 0; k < cnt; k++) {
                const float4* tp = sc.tris + 3 * (int64_t)(first + k);
                float t, u, w;
                if (tri_watertight(pre, tp[0], tp[1], tp[2], tmin, h.t, &t, &u, &w)) {
                    h.t = t;
                    h.tri = first + k;
                    h.u = u;
                    h.v = w;
                }
            }
            leaf = cur;
            if (cur < 0) cur = POP();
        }
        if (cur == EMPTY_REF) break;
    }
#undef POP
    return h;
}

// Geometric normal of a leaf-order triangle (triangle.cc:45-47).
__device__ inline v3f tri_normal(const SceneDev& sc, int32_t t) {
    const float4* tp = sc.tris + 3 * (int64_t)t;
    v3f a = xyz(tp[0]), b = xyz(tp[1]), c = xyz(tp[2]);
    return normalize(cross(b - a, c - a));
}
__device__ inline int32_t tri_orig_id(const SceneDev& sc, int32_t t) {
    return __float_as_int(sc.tris[3 * (int64_t)t].w);
}
