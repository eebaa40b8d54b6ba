// gather.hip -- fixed-radius cone-filtered density estimate (the P-KZ form of
// sources/ppm_probabilistic.cc:300-334) with Kahan accumulation into the image.
//
// For a hit point x with oriented normal n_x, weight W = throughput * rho / pi
// and radius r, every photon p of the hash grid buckets covering the ball with
// |x - p|^2 <= r^2 and n_p . n_x > 0 contributes (1 - |x - p| / (1.1 r)) Phi_p;
// L = W * sum / (pi r^2 (1 - 2 / 3.3)).
//
// Photons are read in bucket (cell) order from a gather layout: for the lanes
// and blocks kernels the records as stored, (x, y, z, octahedral normal)
// (Phi, key), with a box per 32-record block; for the others three 16-byte
// words (x, y, z, n_x) (n_y, n_z, Phi_r, Phi_g) (Phi_b, key, -, -) with the
// normal decoded once.  Hit points arrive sorted along a Morton curve of half
// cells (edge r), so neighbouring lanes share their eight buckets.  Four
// kernels compute the same sum:
//   k_gather_lanes (default): one lane per hit point walking its own buckets;
//     inside a bucket the records follow a cell-local Morton order and blocks
//     of 32 records whose bounding box misses the ball are skipped.
//   k_gather_blocks: a wave walks the union of its lanes' buckets block by
//     block, skips the blocks no lane needs (the lanes kernel's decisions) and
//     stages the others in LDS for all 64 lanes.
//   k_gather_tiled: the union of a wave's buckets is staged through LDS in
//     coalesced 64-photon chunks and tested by all 64 lanes.
//   k_gather_wave: the union is walked with a wave-uniform photon index, so
//     photon records arrive through scalar loads and are tested by all lanes.
#include "internal.h"

namespace {

// Cells of edge 2r met by [x - r, x + r] per axis: at most two (a third can
// only come from FP64 rounding at an exact cell boundary, and is dropped).
__device__ __forceinline__ void hit_cells(const GatherArgs& a, float4 hp, int32_t lo[3], int32_t hi[3]) {
    const double r = gather_radius(a);
    const double sc = grid_scale(a.g);
    const double xs[3] = {hp.x, hp.y, hp.z};
    for (int d = 0; d < 3; d++) {
        lo[d] = grid_cell(xs[d] - r, a.g.bmin[d], sc);
        hi[d] = min(grid_cell(xs[d] + r, a.g.bmin[d], sc), lo[d] + 1);
    }
}

// photon records through the constant address space: loads with a
// wave-uniform index become scalar loads
typedef float f4v __attribute__((ext_vector_type(4)));
typedef const __attribute__((address_space(4))) f4v* const_f4;
__device__ __forceinline__ const_f4 as_const(const float4* p) { return (const_f4)p; }
__device__ __forceinline__ float4 ld(const_f4 p, uint32_t j) {
    const f4v v = p[j];
    return make_float4(v.x, v.y, v.z, v.w);
}

// direction of an octahedral normal code, unnormalised (only its sign
// against n_x is needed)
__device__ __forceinline__ v3f oct_direction(uint32_t e) {
    float qx = (float)(int16_t)(e & 0xFFFF) * (1.0f / 32767.0f);
    float qy = (float)(int16_t)(e >> 16) * (1.0f / 32767.0f);
    const float qz = 1.0f - fabsf(qx) - fabsf(qy);
    if (qz < 0.0f) {
        const float ox = (1.0f - fabsf(qy)) * sgnf(qx);
        const float oy = (1.0f - fabsf(qx)) * sgnf(qy);
        qx = ox;
        qy = oy;
    }
    return mk(qx, qy, qz);
}

// oct_direction(e) . n > 0 in fewer instructions, with the same values: the
// fold's sign factor is a copysign (a component is never -0: int16 0
// converts to +0), and the dot product is contracted in one fixed order.
__device__ __forceinline__ bool normal_agrees(uint32_t e, v3f n) {
    float qx = (float)(int16_t)(e & 0xFFFF) * (1.0f / 32767.0f);
    float qy = (float)(int16_t)(e >> 16) * (1.0f / 32767.0f);
    const float qz = 1.0f - fabsf(qx) - fabsf(qy);
    const float fx = copysignf(1.0f - fabsf(qy), qx), fy = copysignf(1.0f - fabsf(qx), qy);
    qx = qz < 0.0f ? fx : qx;
    qy = qz < 0.0f ? fy : qy;
    return fmaf(qz, n.z, fmaf(qx, n.x, qy * n.y)) > 0.0f;
}

// Smallest bucket id among the slots of mb set in m (all ones if none), and
// the slot set with that bucket cleared: value operations only, so the eight
// slots stay in registers.
__device__ __forceinline__ uint32_t min_bucket(const uint32_t mb[8], uint32_t m) {
    uint32_t v = 0xFFFFFFFFu;
#pragma unroll
    for (int k = 0; k < 8; k++)
        if ((m >> k) & 1u) v = min(v, mb[k]);
    return v;
}
__device__ __forceinline__ uint32_t clear_bucket(const uint32_t mb[8], uint32_t m, uint32_t b) {
#pragma unroll
    for (int k = 0; k < 8; k++)
        if (mb[k] == b) m &= ~(1u << k);
    return m;
}

__device__ __forceinline__ float cone_norm(double r) {
    return (float)(1.0 / (3.14159265358979323846 * r * r * (1.0 - 2.0 / (3.0 * 1.1))));
}

__device__ __forceinline__ void write_result(const GatherArgs& a, int64_t pix, v3f w, float ax, float ay, float az) {
    const float inv_norm = cone_norm(gather_radius(a));
    const float Lx = w.x * ax * inv_norm, Ly = w.y * ay * inv_norm, Lz = w.z * az * inv_norm;
    if (a.img_sum) {
        kahan_add(a.img_sum, a.img_comp, 3 * pix + 0, Lx);
        kahan_add(a.img_sum, a.img_comp, 3 * pix + 1, Ly);
        kahan_add(a.img_sum, a.img_comp, 3 * pix + 2, Lz);
    } else {
        a.out_rgb[3 * pix + 0] = Lx;
        a.out_rgb[3 * pix + 1] = Ly;
        a.out_rgb[3 * pix + 2] = Lz;
    }
}

// The buckets of the (up to 2 x 2 x 2) cells of one hit point in the
// reference's z, y, x loop order; bit c of the returned mask marks slot c as
// a cell of the range whose bucket no earlier slot holds (hash collisions and
// abs() folding can map two cells to one bucket).
__device__ __forceinline__ uint32_t own_buckets(const GatherArgs& a, float4 hp, uint32_t mb[8]) {
    int32_t lo[3], hi[3];
    hit_cells(a, hp, lo, hi);
    uint32_t vm = 0;
#pragma unroll
    for (int c = 0; c < 8; c++) {
        const int dx = c & 1, dy = (c >> 1) & 1, dz = c >> 2;
        const bool in = lo[0] + dx <= hi[0] && lo[1] + dy <= hi[1] && lo[2] + dz <= hi[2];
        mb[c] = grid_hash(lo[0] + dx, lo[1] + dy, lo[2] + dz, a.g.size);
        bool dup = false;
#pragma unroll
        for (int k = 0; k < c; k++) dup |= ((vm >> k) & 1u) && mb[k] == mb[c];
        if (in && !dup) vm |= 1u << c;
    }
    return vm;
}

struct Acc {
    float x = 0.0f, y = 0.0f, z = 0.0f;
    unsigned long long n = 0;
};

// |x - p|^2 with explicit fused operations: every kernel decides the radius
// test of a photon on the same bits
__device__ __forceinline__ float dist2(v3f x, float4 A) {
    const float dx = x.x - A.x, dy = x.y - A.y, dz = x.z - A.z;
    return __fmaf_rn(dx, dx, __fmaf_rn(dy, dy, __fmul_rn(dz, dz)));
}

// cone filter weight 1 - |x - p| / (1.1 r) from the squared distance, with the
// hardware square root (v_sqrt_f32) that every gather kernel shares
__device__ __forceinline__ float cone_weight(float d2, float inv_kr) {
    return 1.0f - __builtin_amdgcn_sqrtf(d2) * inv_kr;
}

// one candidate inside the radius: normal agreement, cone weight
__device__ __forceinline__ void accumulate(Acc& acc, v3f nx, float d2, float inv_kr, float4 A, float4 B, float4 C) {
    if (A.w * nx.x + B.x * nx.y + B.y * nx.z > 0.0f) {
        const float wt = cone_weight(d2, inv_kr);
        acc.x += wt * B.z;
        acc.y += wt * B.w;
        acc.z += wt * C.x;
        acc.n++;
    }
}

__device__ __forceinline__ void flush_counters(const GatherArgs& a, bool valid, unsigned long long nb,
                                      unsigned long long nc, unsigned long long ni,
                                      unsigned long long nt, unsigned long long nk) {
    nb = wave_sum(valid ? nb : 0);
    nc = wave_sum(valid ? nc : 0);
    ni = wave_sum(valid ? ni : 0);
    nt = wave_sum(valid ? nt : 0);
    nk = wave_sum(valid ? nk : 0);
    if ((threadIdx.x & 63) == 0 && a.counters) {
        atomicAdd(&a.counters->gather_buckets, nb);
        atomicAdd(&a.counters->gather_candidates, nc);
        atomicAdd(&a.counters->gather_inradius, ni);
        atomicAdd(&a.counters->gather_tested, nt);
        atomicAdd(&a.counters->gather_blocks, nk);
    }
}

// The records [j0, j1) of one bucket: LANES_INFLIGHT 16-byte records in
// flight per wait (records past the range read and masked out), flux loads
// issued before the sums, accumulation in record order.
__device__ __forceinline__ void lanes_range(const GatherArgs& a, uint32_t j0, uint32_t j1, v3f x, v3f nx, float r2,
                                            float inv_kr, bool test_normals, float& ax, float& ay, float& az,
                                            uint32_t& ni) {
    const float4* __restrict__ pa = a.ph_a;
    const float4* __restrict__ pb = a.ph_b;
    for (uint32_t j = j0; j < j1; j += LANES_INFLIGHT) {
        // up to LANES_INFLIGHT - 1 records past the range are read and masked
        // out: the layout array holds GATHER_PAD records beyond its last one
        const float4* __restrict__ pj = pa + j;
        float4 A[LANES_INFLIGHT];
#pragma unroll
        for (int q = 0; q < LANES_INFLIGHT; q++) A[q] = pj[q];
        // the normal word is read with the position (one 16-byte load)
#pragma unroll
        for (int q = 0; q < LANES_INFLIGHT; q++) asm volatile("" ::"v"(A[q].w));
        float d2[LANES_INFLIGHT];
        bool in[LANES_INFLIGHT];
        // the normal test is skipped where the block's normals agree as one
#pragma unroll
        for (int q = 0; q < LANES_INFLIGHT; q++) {
            d2[q] = dist2(x, A[q]);
            in[q] = j + q < j1 && d2[q] <= r2 && (!test_normals || normal_agrees(__float_as_uint(A[q].w), nx));
        }
        float4 F[LANES_INFLIGHT];
        const float4* __restrict__ pbj = pb + j;
#pragma unroll
        for (int q = 0; q < LANES_INFLIGHT; q++)
            if (in[q]) F[q] = pbj[q];
#pragma unroll
        for (int q = 0; q < LANES_INFLIGHT; q++) {
            if (in[q]) {
                const float wt = cone_weight(d2[q], inv_kr);
                ax += wt * F[q].x;
                ay += wt * F[q].y;
                az += wt * F[q].z;
                ni++;
                if (a.used) a.used[j + q] = 1u;
            }
        }
    }
}

// Squared distance from x to the box (lo, hi), zero inside.
__device__ __forceinline__ float box_dist2(v3f x, float4 lo, float4 hi) {
    const float dx = fmaxf(fmaxf(lo.x - x.x, x.x - hi.x), 0.0f);
    const float dy = fmaxf(fmaxf(lo.y - x.y, x.y - hi.y), 0.0f);
    const float dz = fmaxf(fmaxf(lo.z - x.z, x.z - hi.z), 0.0f);
    return dx * dx + dy * dy + dz * dz;
}

// One lane per hit point over its own buckets.  With block boxes, a bucket is
// walked in blocks of GATHER_BLOCK records and a block whose box lies farther
// than the radius (with a relative margin of 2^-12 on r^2, far above the FP32
// rounding of either distance) holds no accepted record and is skipped; the
// accepted records and their order are those of the full walk.
__global__ __launch_bounds__(256) void k_gather_lanes(GatherArgs a) {
    const int64_t n = a.hp_count_src ? (int64_t)a.hp_count_src->hitpoints : a.n_hit;
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const bool valid = i < n;
    uint32_t nb = 0, nc = 0, nt = 0, nk = 0;
    uint32_t ni = 0;
    float ax = 0.0f, ay = 0.0f, az = 0.0f;
    if (valid) {
        const float4 hp = a.hp_pos[i];
        const v3f x = xyz(hp);
        const v3f nx = oct_decode(__float_as_uint(hp.w));
        // the slots to walk; a slot's bucket id is hashed again when it is
        // walked, and the hit point's weight read at the end, which keeps the
        // registers within eight waves per SIMD
        uint32_t todo;
        int32_t clo[3];
        {
            uint32_t mb[8];
            todo = own_buckets(a, hp, mb);
            int32_t chi[3];
            hit_cells(a, hp, clo, chi);
        }
        const float r2 = (float)(gather_radius(a) * gather_radius(a));
        const float r2_box = r2 * (1.0f + 1.0f / 4096.0f);
        const float inv_kr = (float)(1.0 / (1.1 * gather_radius(a)));
        // slots in the reference's z, y, x cell order: lanes of neighbouring
        // half cells then walk their shared buckets in step
#pragma nounroll
        for (int c = 0; c < 8; c++) {
            if (!((todo >> c) & 1u)) continue;
            const uint32_t bc = grid_hash(clo[0] + (c & 1), clo[1] + ((c >> 1) & 1), clo[2] + (c >> 2), a.g.size);
            const uint32_t b0 = a.off[bc], b1 = a.off[bc + 1];
            nb++;
            nc += b1 - b0;
            if (b1 <= b0) continue;
            // segments: the whole bucket, or its blocks whose box meets the ball
            const bool boxed = a.boxes != nullptr;
            const uint32_t k0 = boxed ? b0 / GATHER_BLOCK : 0, k1 = boxed ? (b1 - 1) / GATHER_BLOCK : 0;
            for (uint32_t k = k0; k <= k1; k++) {
                bool test_normals = true;
                if (boxed) {
                    nk++;
                    const float4 lo = a.boxes[2 * k];
                    if (box_dist2(x, lo, a.boxes[2 * k + 1]) > r2_box) continue;
                    // one normal for the whole block: one test decides for
                    // every record of it
                    const uint32_t eb = __float_as_uint(lo.w);
                    if (eb != MIXED_NORMALS) {
                        if (!normal_agrees(eb, nx)) continue;
                        test_normals = false;
                    }
                }
                const uint32_t j0 = boxed ? max(b0, k * GATHER_BLOCK) : b0;
                const uint32_t j1 = boxed ? min(b1, (k + 1) * GATHER_BLOCK) : b1;
                nt += j1 - j0;
                lanes_range(a, j0, j1, x, nx, r2, inv_kr, test_normals, ax, ay, az, ni);
            }
        }
        const float4 hw = a.hp_w[i];
        write_result(a, (int64_t)__float_as_uint(hw.w), xyz(hw), ax, ay, az);
    }
    flush_counters(a, valid, nb, nc, ni, nt, nk);
}

// Per-lane setup shared by the wave-cooperative kernels.
struct LaneHit {
    bool valid;
    float4 hp, hw;
    v3f x, nx;
    uint32_t mb[8];
    uint32_t vm;  // valid, de-duplicated slots of mb
    unsigned long long nb, nc;
};

__device__ __forceinline__ void lane_hit(const GatherArgs& a, int64_t n, int64_t i, LaneHit& h) {
    h.valid = i < n;
    h.hp = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    h.hw = h.hp;
    if (h.valid) {
        h.hp = a.hp_pos[i];
        h.hw = a.hp_w[i];
    }
    h.x = xyz(h.hp);
    h.nx = h.valid ? oct_decode(__float_as_uint(h.hp.w)) : mk(0.0f, 0.0f, 0.0f);
    const uint32_t vm = own_buckets(a, h.hp, h.mb);  // every lane: no divergent array writes
    h.vm = h.valid ? vm : 0u;
    h.nb = (unsigned long long)__popc(h.vm);
    h.nc = 0;
#pragma unroll
    for (int k = 0; k < 8; k++)
        if ((h.vm >> k) & 1u) h.nc += a.off[h.mb[k] + 1] - a.off[h.mb[k]];
}

// Next bucket of the wave's union (wave-uniform), removing it from every
// lane's pending set; false when the union is exhausted.
// The union is walked in ascending bucket id: the wave minimum of the lanes'
// smallest pending buckets.
template <class Hit>
__device__ __forceinline__ bool next_union_bucket(const Hit& h, uint32_t& pending, uint32_t& bucket) {
    if (__ballot(pending != 0u) == 0) return false;
    uint32_t v = min_bucket(h.mb, pending);
    for (int off = 32; off > 0; off >>= 1) v = min(v, (uint32_t)__shfl_xor(v, off));
    const uint32_t b = __builtin_amdgcn_readfirstlane(v);
    pending = clear_bucket(h.mb, pending, b);
    bucket = b;
    return true;
}

// Four independent waves per workgroup, 64 Morton-adjacent hit points each.
__global__ __launch_bounds__(256) void k_gather_wave(GatherArgs a) {
    const int64_t n = a.hp_count_src ? (int64_t)a.hp_count_src->hitpoints : a.n_hit;
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (((int64_t)blockIdx.x * 256 + (threadIdx.x & ~63)) >= n) return;  // idle wave (uniform)
    LaneHit h;
    lane_hit(a, n, i, h);
    uint32_t pending = h.vm;
    const float r2 = (float)(gather_radius(a) * gather_radius(a));
    const float inv_kr = (float)(1.0 / (1.1 * gather_radius(a)));
    const const_f4 pa = as_const(a.ph_a), pb = as_const(a.ph_b), pc = as_const(a.ph_c);
    const __attribute__((address_space(4))) uint32_t* off =
        (const __attribute__((address_space(4))) uint32_t*)a.off;
    Acc acc;
    uint32_t b;
    while (next_union_bucket(h, pending, b)) {
        const uint32_t j0 = off[b], j1 = off[b + 1];
        uint32_t j = j0;
        // four positions in flight per wait; the bucket's last record repeats
        // past its end and is masked out
        for (; j < j1; j += 4) {
            const uint32_t last = j1 - 1;
            float4 A[4];
#pragma unroll
            for (int q = 0; q < 4; q++) A[q] = ld(pa, min(j + q, last));
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const float d2 = dist2(h.x, A[q]);
                if (j + q < j1 && d2 <= r2) accumulate(acc, h.nx, d2, inv_kr, A[q], ld(pb, j + q), ld(pc, j + q));
            }
        }
    }
    if (h.valid) write_result(a, (int64_t)__float_as_uint(h.hw.w), xyz(h.hw), acc.x, acc.y, acc.z);
    flush_counters(a, h.valid, h.nb, h.nc, acc.n, h.nc, 0);
}

// One wave (64 lanes) per workgroup; the union's photons staged through LDS.
__global__ __launch_bounds__(64) void k_gather_tiled(GatherArgs a) {
    __shared__ float4 s_a[64];
    __shared__ float4 s_b[64];
    __shared__ float4 s_c[64];
    __shared__ uint32_t s_start[64];
    __shared__ uint32_t s_pre[65];
    const int64_t n = a.hp_count_src ? (int64_t)a.hp_count_src->hitpoints : a.n_hit;
    const int lane = threadIdx.x;
    const int64_t i = (int64_t)blockIdx.x * 64 + lane;
    if ((int64_t)blockIdx.x * 64 >= n) return;  // whole workgroup idle (uniform)
    LaneHit h;
    lane_hit(a, n, i, h);
    uint32_t pending = h.vm;
    const float r2 = (float)(gather_radius(a) * gather_radius(a));
    const float inv_kr = (float)(1.0 / (1.1 * gather_radius(a)));
    Acc acc;
    if (lane == 0) s_pre[0] = 0;
    while (true) {
        // up to 64 buckets of the wave's union, one per lane
        uint32_t my_bucket = 0, b;
        int nu = 0;
        while (nu < 64 && next_union_bucket(h, pending, b)) {
            if (lane == nu) my_bucket = b;
            nu++;
        }
        if (nu == 0) break;
        uint32_t start = 0, size = 0;
        if (lane < nu) {
            start = a.off[my_bucket];
            size = a.off[my_bucket + 1] - start;
        }
        uint32_t incl = size;
        for (int off = 1; off < 64; off <<= 1) {
            const uint32_t t = __shfl_up(incl, off);
            if (lane >= off) incl += t;
        }
        const uint32_t total = __shfl(incl, 63);
        __syncthreads();
        s_start[lane] = start;
        s_pre[lane + 1] = incl;
        __syncthreads();
        int u0 = 0;
        for (uint32_t base = 0; base < total; base += 64) {
            const uint32_t g = base + (uint32_t)lane;
            if (g < total) {
                int u = u0;
                while (g >= s_pre[u + 1]) u++;
                const uint32_t j = s_start[u] + (g - s_pre[u]);
                s_a[lane] = a.ph_a[j];
                s_b[lane] = a.ph_b[j];
                s_c[lane] = a.ph_c[j];
            }
            __syncthreads();
            const int m = (total - base) < 64u ? (int)(total - base) : 64;
            for (int k = 0; k < m; k++) {
                const float4 A = s_a[k];
                const float d2 = dist2(h.x, A);
                if (d2 <= r2) accumulate(acc, h.nx, d2, inv_kr, A, s_b[k], s_c[k]);
            }
            __syncthreads();
            const uint32_t next = base + 64;
            while (u0 < nu - 1 && next >= s_pre[u0 + 1]) u0++;
        }
    }
    if (h.valid) write_result(a, (int64_t)__float_as_uint(h.hw.w), xyz(h.hw), acc.x, acc.y, acc.z);
    flush_counters(a, h.valid, h.nb, h.nc, acc.n, h.nc, 0);
}

// Orders a wave's LDS writes before its later LDS reads (the LDS serves one
// wave's operations in order; this keeps the compiler from moving them).
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

#define BLOCKS_UNROLL 4
// The m (<= GATHER_BLOCK) records staged at rec (positions, then fluxes at
// rec + GATHER_BLOCK) for one lane: BLOCKS_UNROLL records per step, read
// before any is used; slots past m hold stale records and are masked out (weight and flux
// zero).  TN: test each record's normal; MARK: used[t] = 1 for accepted t.
template <bool TN, bool MARK>
__device__ __forceinline__ void blocks_walk(const float4* rec, uint32_t m, bool need, v3f x, v3f nx, float r2,
                                            float inv_kr, uint32_t* used, float& ax, float& ay, float& az,
                                            uint32_t& ni) {
    for (uint32_t t = 0; t < m; t += BLOCKS_UNROLL) {
        float4 A[BLOCKS_UNROLL], F[BLOCKS_UNROLL];
#pragma unroll
        for (int u = 0; u < BLOCKS_UNROLL; u++) A[u] = rec[t + u];
#pragma unroll
        for (int u = 0; u < BLOCKS_UNROLL; u++) F[u] = rec[GATHER_BLOCK + t + u];
#pragma unroll
        for (int u = 0; u < BLOCKS_UNROLL; u++) {
            const float d2 = dist2(x, A[u]);
            bool in = need && t + u < m && d2 <= r2;
            if (TN) in = in && normal_agrees(__float_as_uint(A[u].w), nx);
            const float wt = in ? cone_weight(d2, inv_kr) : 0.0f;
            ax = fmaf(wt, in ? F[u].x : 0.0f, ax);
            ay = fmaf(wt, in ? F[u].y : 0.0f, ay);
            az = fmaf(wt, in ? F[u].z : 0.0f, az);
            ni += in ? 1u : 0u;
            if (MARK && in) used[t + u] = 1u;
        }
    }
}

// One wave per 64 Morton-adjacent hit points, four independent waves per
// workgroup.  The wave walks the union of its lanes' buckets block by block
// (GATHER_BLOCK records): a lane needs a block when the block's bucket is one
// of its own, the block's box meets its ball and, for a block of one normal
// code, that code agrees with its normal -- the lanes kernel's decisions, so
// each lane accepts the same records.  A block no lane needs is skipped by
// the whole wave; otherwise its records are staged in LDS (position words by
// lanes 0-31, flux words by lanes 32-63, one coalesced load) and every lane
// walks them in step, reading each record by broadcast.  Per lane the sum
// runs over its accepted records in union order (ascending bucket id).
// Squared distance between the boxes (alo, ahi) and (lo, hi), zero where they
// meet: a lower bound, in FP32 too, of box_dist2(x, lo, hi) for x in the first.
__device__ __forceinline__ float box_box_dist2(const float alo[3], const float ahi[3], float4 lo, float4 hi) {
    const float dx = fmaxf(fmaxf(lo.x - ahi[0], alo[0] - hi.x), 0.0f);
    const float dy = fmaxf(fmaxf(lo.y - ahi[1], alo[1] - hi.y), 0.0f);
    const float dz = fmaxf(fmaxf(lo.z - ahi[2], alo[2] - hi.z), 0.0f);
    return dx * dx + dy * dy + dz * dz;
}

// A lane's hit point in the blocks kernel: the fields next_union_bucket reads
struct BlockHit {
    bool valid;
    v3f x, nx;
    uint32_t mb[8];
    uint32_t vm;  // valid, de-duplicated slots of mb
};

// Stages the m records [j0, j0 + m) of one block (positions by lanes 0-31,
// fluxes by lanes 32-63, one coalesced load) and walks them for this lane.
template <bool MARK>
__device__ __forceinline__ void blocks_stage_walk(const GatherArgs& a, float4* rec, int lane, uint32_t j0, uint32_t m,
                                                  bool need, bool test_normals, const BlockHit& h, float r2,
                                                  float inv_kr, float& ax, float& ay, float& az, uint32_t& ni) {
    const uint32_t q = (uint32_t)lane & (GATHER_BLOCK - 1);
    if (q < m) rec[lane] = lane < GATHER_BLOCK ? a.ph_a[j0 + q] : a.ph_b[j0 + q];
    wave_lds_sync();
    if (test_normals)
        blocks_walk<true, MARK>(rec, m, need, h.x, h.nx, r2, inv_kr, a.used + j0, ax, ay, az, ni);
    else
        blocks_walk<false, MARK>(rec, m, need, h.x, h.nx, r2, inv_kr, a.used + j0, ax, ay, az, ni);
    wave_lds_sync();  // every lane has read the block before the next one is staged
}

template <bool MARK>
__global__ __launch_bounds__(256) void k_gather_blocks(GatherArgs a) {
    __shared__ float4 s_rec[4][2 * GATHER_BLOCK];
    __shared__ float4 s_box[4][2 * 64];
    const int lane = threadIdx.x & 63;
    const int64_t n = a.hp_count_src ? (int64_t)a.hp_count_src->hitpoints : a.n_hit;
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (((int64_t)blockIdx.x * 256 + (threadIdx.x & ~63)) >= n) return;  // idle wave (uniform)
    float4* rec = s_rec[threadIdx.x >> 6];
    float4* sbox = s_box[threadIdx.x >> 6];
    // the lane's hit point (its weight is read at the end)
    BlockHit h;
    h.valid = i < n;
    const float4 hp = h.valid ? a.hp_pos[i] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    h.x = xyz(hp);
    h.nx = h.valid ? oct_decode(__float_as_uint(hp.w)) : mk(0.0f, 0.0f, 0.0f);
    {
        const uint32_t vm = own_buckets(a, hp, h.mb);
        h.vm = h.valid ? vm : 0u;
    }
    uint32_t nb = (uint32_t)__popc(h.vm), nc = 0;
#pragma unroll
    for (int k = 0; k < 8; k++)
        if ((h.vm >> k) & 1u) nc += a.off[h.mb[k] + 1] - a.off[h.mb[k]];
    uint32_t pending = h.vm;
    const float r2 = (float)(gather_radius(a) * gather_radius(a));
    const float r2_box = r2 * (1.0f + 1.0f / 4096.0f);
    const float inv_kr = (float)(1.0 / (1.1 * gather_radius(a)));
    // the box of the wave's hit points: a block farther from it than the
    // radius is needed by no lane
    float wl[3] = {h.valid ? h.x.x : 3.0e38f, h.valid ? h.x.y : 3.0e38f, h.valid ? h.x.z : 3.0e38f};
    float wh[3] = {h.valid ? h.x.x : -3.0e38f, h.valid ? h.x.y : -3.0e38f, h.valid ? h.x.z : -3.0e38f};
#pragma unroll
    for (int off = 32; off > 0; off >>= 1)
#pragma unroll
        for (int d = 0; d < 3; d++) {
            wl[d] = fminf(wl[d], __shfl_xor(wl[d], off));
            wh[d] = fmaxf(wh[d], __shfl_xor(wh[d], off));
        }
    float ax = 0.0f, ay = 0.0f, az = 0.0f;
    uint32_t ni = 0, nt = 0, nk = 0, staged = 0;
    uint32_t b;
    while (next_union_bucket(h, pending, b)) {
        bool own = false;
#pragma unroll
        for (int c = 0; c < 8; c++) own |= ((h.vm >> c) & 1u) && h.mb[c] == b;
        const uint32_t b0 = a.off[b], b1 = a.off[b + 1];
        if (b1 <= b0) continue;
        const uint32_t k0 = b0 / GATHER_BLOCK, k1 = (b1 - 1) / GATHER_BLOCK;
        if (!a.boxes) {  // no block boxes: every block of the bucket, no culling
            for (uint32_t k = k0; k <= k1; k++) {
                const uint32_t j0 = max(b0, k * GATHER_BLOCK), m = min(b1, (k + 1) * GATHER_BLOCK) - j0;
                staged += m;
                nt += own ? m : 0u;
                blocks_stage_walk<MARK>(a, rec, lane, j0, m, own, true, h, r2, inv_kr, ax, ay, az, ni);
            }
            continue;
        }
        nk += own ? k1 - k0 + 1 : 0u;
        // 64 block boxes per step, one per lane: a coalesced load into LDS
        // and the test against the wave's box; then the candidates one by one
        for (uint32_t kc = k0; kc <= k1; kc += 64) {
            const uint32_t kk = kc + (uint32_t)lane;
            bool cand = false;
            if (kk <= k1) {
                const float4 lo = a.boxes[2 * kk], hi = a.boxes[2 * kk + 1];
                sbox[2 * lane] = lo;
                sbox[2 * lane + 1] = hi;
                cand = box_box_dist2(wl, wh, lo, hi) <= r2_box;
            }
            unsigned long long cm = __ballot(cand);
            wave_lds_sync();
            while (cm) {
                const int t = __builtin_ctzll(cm);
                cm &= cm - 1;
                const float4 lo = sbox[2 * t];
                bool need = own && box_dist2(h.x, lo, sbox[2 * t + 1]) <= r2_box;
                bool test_normals = true;
                const uint32_t eb = __float_as_uint(lo.w);
                if (eb != MIXED_NORMALS) {  // one normal code for the whole block
                    need = need && normal_agrees(eb, h.nx);
                    test_normals = false;
                }
                if (__ballot(need) == 0) continue;  // no lane's ball meets the block
                const uint32_t k = kc + (uint32_t)t;
                const uint32_t j0 = max(b0, k * GATHER_BLOCK), m = min(b1, (k + 1) * GATHER_BLOCK) - j0;
                staged += m;
                nt += need ? m : 0u;
                blocks_stage_walk<MARK>(a, rec, lane, j0, m, need, test_normals, h, r2, inv_kr, ax, ay, az, ni);
            }
            wave_lds_sync();  // the boxes are read before the next step overwrites them
        }
    }
    if (h.valid) {
        const float4 hw = a.hp_w[i];
        write_result(a, (int64_t)__float_as_uint(hw.w), xyz(hw), ax, ay, az);
    }
    flush_counters(a, h.valid, nb, nc, ni, nt, nk);
    if (lane == 0 && a.counters) atomicAdd(&a.counters->gather_staged, (unsigned long long)staged);
}

// The box and shared normal code of the block of GATHER_BLOCK records that
// record j belongs to, from record p on each of the block's 32 consecutive
// lanes: half-wave min / max reductions; the block's first lane writes them.
__device__ __forceinline__ void block_box(float4 p, int64_t j, float4* __restrict__ boxes) {
    const int64_t k = j / GATHER_BLOCK;
    float lx = p.x, ly = p.y, lz = p.z, hx = p.x, hy = p.y, hz = p.z;
    uint32_t emin = __float_as_uint(p.w), emax = emin;
#pragma unroll
    for (int off = GATHER_BLOCK / 2; off > 0; off >>= 1) {
        lx = fminf(lx, __shfl_xor(lx, off));
        ly = fminf(ly, __shfl_xor(ly, off));
        lz = fminf(lz, __shfl_xor(lz, off));
        hx = fmaxf(hx, __shfl_xor(hx, off));
        hy = fmaxf(hy, __shfl_xor(hy, off));
        hz = fmaxf(hz, __shfl_xor(hz, off));
        emin = min(emin, (uint32_t)__shfl_xor(emin, off));
        emax = max(emax, (uint32_t)__shfl_xor(emax, off));
    }
    if ((j % GATHER_BLOCK) == 0) {
        const uint32_t e = emin == emax ? emin : MIXED_NORMALS;
        boxes[2 * k] = make_float4(lx, ly, lz, __uint_as_float(e));
        boxes[2 * k + 1] = make_float4(hx, hy, hz, 0.0f);
    }
}

// Box of each block of GATHER_BLOCK consecutive records (GATHER_BLOCK = 32
// lanes per block).
__global__ __launch_bounds__(256) void k_gather_boxes(const float4* __restrict__ a, int64_t n,
                                                      float4* __restrict__ boxes) {
    const int64_t j = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const int64_t nbox = (n + GATHER_BLOCK - 1) / GATHER_BLOCK;
    if (j / GATHER_BLOCK >= nbox) return;  // whole blocks of 32 lanes leave together
    block_box(a[min(j, n - 1)], j, boxes);  // the last record repeated past the end
}

__device__ __forceinline__ uint32_t spread10(uint32_t v) {
    v = (v * 0x00010001u) & 0xFF0000FFu;
    v = (v * 0x00000101u) & 0x0F00F00Fu;
    v = (v * 0x00000011u) & 0xC30C30C3u;
    v = (v * 0x00000005u) & 0x49249249u;
    return v;
}

// Morton code of the half cell (edge r) of each hit point: hit points with the
// same code share their eight buckets exactly.
__global__ void k_hit_keys(const float4* __restrict__ hp_pos, const float4* __restrict__ hp_w,
                           int64_t cap, GridParams g, double radius, uint32_t* keys, uint32_t* vals) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= cap) return;
    if (__float_as_uint(hp_w[i].w) == NO_HITPOINT) {
        keys[i] = 0xFFFFFFFFu;
        vals[i] = (uint32_t)i;
        return;
    }
    const float4 p = hp_pos[i];
    const double inv = 1.0 / (g.pd ? g.pd->radius : radius);
    const double c[3] = {((double)p.x - g.bmin[0]) * inv, ((double)p.y - g.bmin[1]) * inv,
                         ((double)p.z - g.bmin[2]) * inv};
    uint32_t q[3];
    for (int d = 0; d < 3; d++) q[d] = (uint32_t)fmin(fmax(c[d], 0.0), 1023.0);
    keys[i] = (spread10(q[0]) << 2) | (spread10(q[1]) << 1) | spread10(q[2]);
    vals[i] = (uint32_t)i;
}

// Compact gather layout with the block boxes of the lanes kernel in the same
// pass: record k = photon perm[sorted[k]]; past n, lanes take record n - 1 for
// the box of the last block and write nothing else.
__global__ __launch_bounds__(256) void k_gather_layout_boxes(const float4* __restrict__ pos,
                                                             const float4* __restrict__ flux,
                                                             const uint32_t* __restrict__ perm,
                                                             const uint32_t* __restrict__ sorted, int64_t n,
                                                             float4* __restrict__ A, float4* __restrict__ B,
                                                             float4* __restrict__ boxes, int stride) {
    const int64_t k = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const int64_t nbox = (n + GATHER_BLOCK - 1) / GATHER_BLOCK;
    if (k / GATHER_BLOCK >= nbox) return;  // whole blocks of 32 lanes leave together
    const int64_t kk = min(k, n - 1);
    const uint32_t j = sorted ? sorted[kk] : (uint32_t)kk;
    const uint32_t src = perm ? perm[j] : j;
    const float4 p = pos[(int64_t)src * stride];
    if (k < n) {
        A[k] = p;
        B[k] = flux[(int64_t)src * stride];
    }
    block_box(p, k, boxes);
}

// gather layout of the photons in bucket order: record k = photon perm[sorted[k]]
__global__ void k_gather_layout(const float4* __restrict__ pos, const float4* __restrict__ flux,
                                const uint32_t* __restrict__ perm, const uint32_t* __restrict__ sorted,
                                int64_t n, float4* __restrict__ A, float4* __restrict__ B,
                                float4* __restrict__ Cc, int stride) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const uint32_t j = sorted ? sorted[k] : (uint32_t)k;
    const uint32_t src = perm ? perm[j] : j;
    const float4 p = pos[(int64_t)src * stride];
    const float4 f = flux[(int64_t)src * stride];
    if (!Cc) {  // compact: the records as stored
        A[k] = p;
        B[k] = f;
        return;
    }
    const v3f nrm = oct_decode(__float_as_uint(p.w));
    A[k] = make_float4(p.x, p.y, p.z, nrm.x);
    B[k] = make_float4(nrm.y, nrm.z, f.x, f.y);
    Cc[k] = make_float4(f.z, f.w, 0.0f, 0.0f);
}

// used[sorted[j]] = 1 for every photon record j (bucket order) that some hit
// point's gather accepts: inside the radius with agreeing normals, the same
// tests as the gather kernels
__global__ __launch_bounds__(256) void k_mark_used(GatherArgs a, const uint32_t* sorted, uint32_t* used) {
    const int64_t n = a.hp_count_src ? (int64_t)a.hp_count_src->hitpoints : a.n_hit;
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const float4 hp = a.hp_pos[i];
    const v3f x = xyz(hp);
    const v3f nx = oct_decode(__float_as_uint(hp.w));
    const float r2 = (float)(gather_radius(a) * gather_radius(a));
    const bool compact = gather_compact(a.mode);
    uint32_t mb[8];
    const uint32_t todo = own_buckets(a, hp, mb);
    for (int c = 0; c < 8; c++) {
        if (!((todo >> c) & 1u)) continue;
        const uint32_t b0 = a.off[mb[c]], b1 = a.off[mb[c] + 1];
        for (uint32_t j = b0; j < b1; j++) {
            const float4 A = a.ph_a[j];
            if (dist2(x, A) > r2) continue;
            v3f np;
            if (compact) {
                np = oct_direction(__float_as_uint(A.w));
            } else {
                const float4 B = a.ph_b[j];
                np = mk(A.w, B.x, B.y);
            }
            if (np.x * nx.x + np.y * nx.y + np.z * nx.z > 0.0f) used[sorted ? sorted[j] : j] = 1u;
        }
    }
}

// used_order[sorted[j]] = used[j]: bucket order -> the order of `sorted`'s values
__global__ __launch_bounds__(256) void k_scatter_used(const uint32_t* used, const uint32_t* sorted, int64_t n,
                                                      uint32_t* used_order) {
    const int64_t j = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (j < n) used_order[sorted[j]] = used[j];
}

}  // namespace

void launch_scatter_used(hipStream_t s, const uint32_t* used, const uint32_t* sorted, int64_t n, uint32_t* used_order) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_scatter_used, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, used, sorted, n, used_order);
}

void launch_mark_used(hipStream_t s, const GatherArgs& a, const uint32_t* sorted, uint32_t* used) {
    if (a.n_hit <= 0) return;
    hipLaunchKernelGGL(k_mark_used, dim3((unsigned)((a.n_hit + 255) / 256)), dim3(256), 0, s, a, sorted, used);
}

static inline unsigned grid1(int64_t n, int b) { return (unsigned)((n + b - 1) / b); }

void launch_gather(hipStream_t s, const GatherArgs& a) {
    // with a device-side count the grid covers the worst case (one per pixel)
    if (a.n_hit <= 0) return;
    if (a.mode == GATHER_LANES)
        hipLaunchKernelGGL(k_gather_lanes, dim3(grid1(a.n_hit, 256)), dim3(256), 0, s, a);
    else if (a.mode == GATHER_BLOCKS && a.used)
        hipLaunchKernelGGL(k_gather_blocks<true>, dim3(grid1(a.n_hit, 256)), dim3(256), 0, s, a);
    else if (a.mode == GATHER_BLOCKS)
        hipLaunchKernelGGL(k_gather_blocks<false>, dim3(grid1(a.n_hit, 256)), dim3(256), 0, s, a);
    else if (a.mode == GATHER_TILED)
        hipLaunchKernelGGL(k_gather_tiled, dim3(grid1(a.n_hit, 64)), dim3(64), 0, s, a);
    else
        hipLaunchKernelGGL(k_gather_wave, dim3(grid1(a.n_hit, 256)), dim3(256), 0, s, a);
}

void launch_gather_boxes(hipStream_t s, const float4* a, int64_t n, float4* boxes) {
    if (n <= 0) return;
    const int64_t threads = gather_box_count(n) * GATHER_BLOCK;
    hipLaunchKernelGGL(k_gather_boxes, dim3(grid1(threads, 256)), dim3(256), 0, s, a, n, boxes);
}

void launch_hit_keys(hipStream_t s, const float4* hp_pos, const float4* hp_w, int64_t cap, GridParams g,
                     double radius, uint32_t* keys, uint32_t* vals) {
    if (cap <= 0) return;
    hipLaunchKernelGGL(k_hit_keys, dim3(grid1(cap, 256)), dim3(256), 0, s, hp_pos, hp_w, cap, g, radius, keys,
                       vals);
}

void launch_gather_layout(hipStream_t s, const float4* pos, const float4* flux, const uint32_t* perm,
                          const uint32_t* sorted, int64_t n, float4* A, float4* B, float4* C, float4* boxes,
                          int stride) {
    if (n <= 0) return;
    if (boxes && !C) {
        const int64_t threads = gather_box_count(n) * GATHER_BLOCK;
        hipLaunchKernelGGL(k_gather_layout_boxes, dim3(grid1(threads, 256)), dim3(256), 0, s, pos, flux, perm,
                           sorted, n, A, B, boxes, stride);
        return;
    }
    hipLaunchKernelGGL(k_gather_layout, dim3(grid1(n, 256)), dim3(256), 0, s, pos, flux, perm, sorted, n, A, B,
                       C, stride);
}
