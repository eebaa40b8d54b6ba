// traverse.h -- closest-hit traversal of the 4-wide BVH (device code).
//
// One ray per lane.  The traversal stack keeps its top STACK_LDS entries in
// LDS (laid out [slot][lane] so the 64 lanes of a wave hit 64 different banks)
// and spills deeper entries to a per-lane private array.  Nodes are fetched as
// seven 16-byte buffer loads from one 128-byte line: per axis the near plane
// (lo for a non-negative direction component, else hi) and the far plane, so
// the slab test needs no min / max to order them; children are visited
// near-first.
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

typedef uint32_t v4u __attribute__((ext_vector_type(4)));

// Word 3 of a raw buffer resource on CDNA: 32-bit data format, no swizzle.
#define BUFFER_RSRC_WORD3 0x00020000

// Pop the traversal stack; EMPTY_REF when it is empty.  The LDS slot is read
// at a clamped index so that only the rare private-array case branches.
template <int BLOCK>
__device__ inline int32_t stack_pop(int& sp, const int32_t* __restrict__ stk, const int32_t* __restrict__ ovf) {
    const int s = sp - 1;
    int32_t v = stk[min(max(s, 0), STACK_LDS - 1) * BLOCK];
    // keeps the two reads apart: merged, they become one flat load of a
    // selected LDS-or-private address, slower than an LDS read
    asm volatile("" : "+v"(v));
    if (s >= STACK_LDS) v = ovf[s - STACK_LDS];
    sp = max(s, 0);
    return s < 0 ? (int32_t)EMPTY_REF : v;
}

// A ray set up for traversal: the watertight test's shear and the slab
// test's scaled inverse direction.
struct TraceRay {
    RayPre pre;  // pre.o is the origin
    float ix, iy, iz, ox, oy, oz;
    float tmin;
    uint32_t near_off;  // byte offsets of the near x / y / z planes in a node, 8 bits each
};

// The resumable part of a traversal.  Between rounds cur is the next inner
// node or EMPTY_REF once the ray is finished (then h is its closest hit); a
// leaf root is kept in cur until the first round tests it.
struct TraceState {
    int32_t cur;
    int32_t sp;
    Hit h;
};

__device__ inline void trace_start(const SceneDev& sc, v3f o, v3f d, float tmin, float tmax, TraceRay& r,
                                   TraceState& s) {
    r.pre = ray_pre(o, d);
    // inverse direction as the reference computes it (qbvh_accel.cc:219-221)
    r.ix = d.x == 0.0f ? 1.0e20f : 1.0f / d.x;
    r.iy = d.y == 0.0f ? 1.0e20f : 1.0f / d.y;
    r.iz = d.z == 0.0f ? 1.0e20f : 1.0f / d.z;
    r.ox = o.x * r.ix;
    r.oy = o.y * r.iy;
    r.oz = o.z * r.iz;
    r.tmin = tmin;
    // Node4: lo x at byte 0, hi x at 16, lo y 32, hi y 48, lo z 64, hi z 80
    r.near_off = (r.ix >= 0.0f ? 0u : 16u) | ((r.iy >= 0.0f ? 32u : 48u) << 8) | ((r.iz >= 0.0f ? 64u : 80u) << 16);
    s.cur = sc.root;
    s.sp = 0;
    s.h.t = tmax;
    s.h.tri = -1;
    s.h.u = 0.0f;
    s.h.v = 0.0f;
}

// Runs rounds of the while-while loop (inner nodes until every busy lane has
// a leaf waiting, then the waiting leaves) over the lanes whose ray is not
// finished, and returns once at most `keep` lanes of the wave are unfinished
// (keep = 0: every ray finished).  Lanes with s.cur == EMPTY_REF on entry take
// no part.  The stack lives in stk (LDS, stride BLOCK) and ovf (private).
template <int BLOCK, bool COUNT = false>
__device__ inline void trace_resume(const SceneDev& sc, const TraceRay& r, TraceState& s,
                                    int32_t* __restrict__ stk, int32_t* __restrict__ ovf, bool* overflow,
                                    TravCount* tc, int keep) {
    const float ix = r.ix, iy = r.iy, iz = r.iz, ox = r.ox, oy = r.oy, oz = r.oz, tmin = r.tmin;
    int32_t cur = s.cur;
    int sp = s.sp;
    Hit h = s.h;
    // cur: an inner node (>= 0), a leaf (< 0) or EMPTY_REF (the stack is
    // empty); leaf: a leaf waiting for its triangle tests (< 0) or none
#define POP() stack_pop<BLOCK>(sp, stk, ovf)
    // the node array as a buffer: 32-bit per-lane byte offsets, and a fetch
    // past its end reads zeros instead of faulting
    const __amdgpu_buffer_rsrc_t nbuf = __builtin_amdgcn_make_buffer_rsrc(
        (void*)sc.nodes, (short)0, sc.n_nodes * (int32_t)sizeof(Node4), BUFFER_RSRC_WORD3);
    const uint32_t nxo = r.near_off & 0xFFu, nyo = (r.near_off >> 8) & 0xFFu, nzo = r.near_off >> 16;
    while (true) {
        int32_t leaf = EMPTY_REF;
        if (cur < 0) {  // a leaf root
            leaf = cur;
            cur = POP();
        }
        // inner nodes, until every lane of the wave has a leaf waiting
        while ((uint32_t)cur < (uint32_t)EMPTY_REF) {
            if (COUNT) tc->nodes++;
            const uint32_t base = (uint32_t)cur << 7;
            const uint32_t ax = base | nxo, ay = base | nyo, az = base | nzo;
            const v4u nx4 = __builtin_amdgcn_raw_buffer_load_b128(nbuf, (int)ax, 0, 0);
            const v4u fx4 = __builtin_amdgcn_raw_buffer_load_b128(nbuf, (int)(ax ^ 16u), 0, 0);
            const v4u ny4 = __builtin_amdgcn_raw_buffer_load_b128(nbuf, (int)ay, 0, 0);
            const v4u fy4 = __builtin_amdgcn_raw_buffer_load_b128(nbuf, (int)(ay ^ 16u), 0, 0);
            const v4u nz4 = __builtin_amdgcn_raw_buffer_load_b128(nbuf, (int)az, 0, 0);
            const v4u fz4 = __builtin_amdgcn_raw_buffer_load_b128(nbuf, (int)(az ^ 16u), 0, 0);
            const v4u ch = __builtin_amdgcn_raw_buffer_load_b128(nbuf, (int)base, 96, 0);
            // one 32-bit key per child: the entry distance (non-negative, so
            // its bit pattern orders like the value) with the child slot in the
            // two low mantissa bits; a missed child keys above every distance.
            // With the planes ordered by the direction's signs these are the
            // min / max of the two slab distances per axis, bit for bit
            // (rounding is monotonic), and an empty slot (every plane +inf)
            // still misses.
            uint32_t key[4];
#pragma unroll
            for (int c = 0; c < 4; c++) {
                const float x0 = __uint_as_float(nx4[c]) * ix - ox, x1 = __uint_as_float(fx4[c]) * ix - ox;
                const float y0 = __uint_as_float(ny4[c]) * iy - oy, y1 = __uint_as_float(fy4[c]) * iy - oy;
                const float z0 = __uint_as_float(nz4[c]) * iz - oz, z1 = __uint_as_float(fz4[c]) * iz - oz;
                const float n = fmaxf(fmaxf(x0, y0), fmaxf(z0, tmin));
                const float f = fminf(fminf(x1, y1), fminf(z1, h.t));
                key[c] = (n <= f * SLAB_FAR_SCALE) ? ((__float_as_uint(n) & 0x7FFFFFFCu) | (uint32_t)c)
                                                   : (0xFFFFFFFCu | (uint32_t)c);
            }
            // sorting network, ascending (near child first), on (key, ref)
            // pairs: each comparator moves the refs with the keys, so no ref
            // is looked up by slot afterwards (keys are distinct: the slot
            // bits break ties)
            uint32_t k0 = key[0], k1 = key[1], k2 = key[2], k3 = key[3];
            int32_t r0 = (int32_t)ch[0], r1 = (int32_t)ch[1], r2 = (int32_t)ch[2], r3 = (int32_t)ch[3];
    #define PSORT(ka, kb, ra, rb)               \
        {                                       \
            const bool sw = kb < ka;            \
            const uint32_t kl = sw ? kb : ka;   \
            kb = sw ? ka : kb;                  \
            ka = kl;                            \
            const int32_t rl = sw ? rb : ra;    \
            rb = sw ? ra : rb;                  \
            ra = rl;                            \
        }
            PSORT(k0, k1, r0, r1);
            PSORT(k2, k3, r2, r3);
            PSORT(k0, k2, r0, r2);
            PSORT(k1, k3, r1, r3);
            PSORT(k1, k2, r1, r2);
    #undef PSORT
            const uint32_t MISS = 0x7F800000u;
            int32_t push[3] = {r3, r2, r1};
            bool pv[3] = {k3 < MISS, k2 < MISS, k1 < MISS};
            if (sp <= STACK_LDS - 3) {
                // room for three in LDS: every candidate is written to the
                // slot above the top, which only a valid one claims
#pragma unroll
                for (int q = 0; q < 3; q++) {
                    stk[sp * BLOCK] = push[q];
                    sp += pv[q] ? 1 : 0;
                }
            } else {
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
            }
            cur = k0 < MISS ? r0 : POP();
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
                if (tri_watertight(r.pre, tp[0], tp[1], tp[2], tmin, h.t, &t, &u, &w)) {
                    h.t = t;
                    h.tri = first + k;
                    h.u = u;
                    h.v = w;
                }
            }
            leaf = cur;
            if (cur < 0) cur = POP();
        }
        if (__popcll(__ballot(cur != (int32_t)EMPTY_REF)) <= keep) break;
    }
#undef POP
    s.cur = cur;
    s.sp = sp;
    s.h = h;
}

// Closest hit of one ray per lane, every lane's traversal run to its end.
template <int BLOCK, bool COUNT = false>
__device__ inline Hit trace_closest(const SceneDev& sc, v3f o, v3f d, float tmin, float tmax,
                                    int32_t* __restrict__ stk, bool* overflow, TravCount* tc = nullptr) {
    TraceRay r;
    TraceState s;
    trace_start(sc, o, d, tmin, tmax, r, s);
    int32_t ovf[STACK_OVF];
    trace_resume<BLOCK, COUNT>(sc, r, s, stk, ovf, overflow, tc, 0);
    return s.h;
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
