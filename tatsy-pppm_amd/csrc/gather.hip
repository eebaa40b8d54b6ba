// gather.hip -- fixed-radius cone-filtered density estimate (the P-KZ form of
// sources/ppm_probabilistic.cc:300-334) with Kahan accumulation into the image.
//
// For a hit point x with oriented normal n_x, weight W = throughput * rho / pi
// and radius r, every photon p of the hash grid buckets covering the ball with
// |x - p|^2 <= r^2 and n_p . n_x > 0 contributes (1 - |x - p| / (1.1 r)) Phi_p;
// L = W * sum / (pi r^2 (1 - 2 / 3.3)).
//
// Two kernels compute the same sum:
//   k_gather_tiled (default): hit points sorted along a Morton curve of half
//     cells; each 64-lane workgroup forms the union of its lanes' buckets with
//     ballots, streams the union's photons through LDS in coalesced chunks of
//     64, and every lane tests its hit point against every staged photon.
//   k_gather_lanes: one lane per hit point walking its own buckets.
#include "internal.h"

namespace {

#define CONE_K 1.1f

struct HitCtx {
    v3f x, nx, w;
    int64_t pix;
    int32_t lo[3], hi[3];
};

__device__ inline void hit_cells(const GatherArgs& a, float4 hp, int32_t lo[3], int32_t hi[3]) {
    const double r = a.radius;
    const double xs[3] = {hp.x, hp.y, hp.z};
    for (int d = 0; d < 3; d++) {
        lo[d] = grid_cell(xs[d] - r, a.g.bmin[d], a.g.scale);
        hi[d] = grid_cell(xs[d] + r, a.g.bmin[d], a.g.scale);
    }
}

// sign of n_p . n_x needs no normalisation of the decoded octahedral n_p
__device__ inline v3f oct_direction(uint32_t e) {
    float qx = (float)(int16_t)(e & 0xFFFF) * (1.0f / 32767.0f);
    float qy = (float)(int16_t)(e >> 16) * (1.0f / 32767.0f);
    float qz = 1.0f - fabsf(qx) - fabsf(qy);
    if (qz < 0.0f) {
        const float ox = (1.0f - fabsf(qy)) * sgnf(qx);
        const float oy = (1.0f - fabsf(qx)) * sgnf(qy);
        qx = ox;
        qy = oy;
    }
    return mk(qx, qy, qz);
}

__device__ inline float cone_norm(double r) {
    return (float)(1.0 / (3.14159265358979323846 * r * r * (1.0 - 2.0 / (3.0 * 1.1))));
}

__device__ inline void write_result(const GatherArgs& a, int64_t i, int64_t pix, v3f w, float ax, float ay,
                                    float az) {
    const float inv_norm = cone_norm(a.radius);
    const float Lx = w.x * ax * inv_norm, Ly = w.y * ay * inv_norm, Lz = w.z * az * inv_norm;
    if (a.img_sum) {
        kahan_add(a.img_sum, a.img_comp, 3 * pix + 0, Lx);
        kahan_add(a.img_sum, a.img_comp, 3 * pix + 1, Ly);
        kahan_add(a.img_sum, a.img_comp, 3 * pix + 2, Lz);
    } else {
        (void)i;
        a.out_rgb[3 * pix + 0] = Lx;
        a.out_rgb[3 * pix + 1] = Ly;
        a.out_rgb[3 * pix + 2] = Lz;
    }
}

// the de-duplicated buckets of one hit point (at most 2 cells per axis: cell edge 2r)
__device__ inline int own_buckets(const GatherArgs& a, float4 hp, uint32_t mb[8]) {
    int32_t lo[3], hi[3];
    hit_cells(a, hp, lo, hi);
    int ns = 0;
    for (int32_t iz = lo[2]; iz <= hi[2]; iz++)
        for (int32_t iy = lo[1]; iy <= hi[1]; iy++)
            for (int32_t ix = lo[0]; ix <= hi[0]; ix++) {
                const uint32_t h = grid_hash(ix, iy, iz, a.g.size);
                bool dup = false;
                for (int k = 0; k < ns; k++) dup |= (mb[k] == h);
                if (!dup && ns < 8) mb[ns++] = h;
            }
    return ns;
}

__global__ __launch_bounds__(256) void k_gather_lanes(GatherArgs a) {
    const int64_t n = a.hp_count_src ? (int64_t)a.hp_count_src->hitpoints : a.n_hit;
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    unsigned long long nb = 0, nc = 0, ni = 0;
    if (i < n) {
        const float4 hp = a.hp_pos[i];
        const float4 hw = a.hp_w[i];
        const v3f x = xyz(hp);
        const v3f nx = oct_decode(__float_as_uint(hp.w));
        uint32_t mb[8];
        const int ns = own_buckets(a, hp, mb);
        const float r2 = (float)(a.radius * a.radius);
        const float inv_kr = (float)(1.0 / (1.1 * a.radius));
        float ax = 0.0f, ay = 0.0f, az = 0.0f;
        for (int k = 0; k < ns; k++) {
            const uint32_t h = mb[k];
            const uint32_t b0 = a.off[h], b1 = a.off[h + 1];
            nb++;
            nc += b1 - b0;
            for (uint32_t j = b0; j < b1; j++) {
                const float4 pp = a.ph_pos[j];
                const float dx = x.x - pp.x, dy = x.y - pp.y, dz = x.z - pp.z;
                const float d2 = dx * dx + dy * dy + dz * dz;
                if (d2 > r2) continue;
                if (dot(oct_direction(__float_as_uint(pp.w)), nx) <= 0.0f) continue;
                ni++;
                const float4 f = a.ph_flux[j];
                const float wt = 1.0f - sqrtf(d2) * inv_kr;
                ax += wt * f.x;
                ay += wt * f.y;
                az += wt * f.z;
            }
        }
        write_result(a, i, (int64_t)__float_as_uint(hw.w), xyz(hw), ax, ay, az);
    }
    nb = wave_sum(nb);
    nc = wave_sum(nc);
    ni = wave_sum(ni);
    if ((threadIdx.x & 63) == 0 && a.counters) {
        atomicAdd(&a.counters->gather_buckets, nb);
        atomicAdd(&a.counters->gather_candidates, nc);
        atomicAdd(&a.counters->gather_inradius, ni);
    }
}

// One wave (64 lanes) per workgroup, 64 Morton-adjacent hit points.
__global__ __launch_bounds__(64) void k_gather_tiled(GatherArgs a) {
    __shared__ float4 s_pos[64];
    __shared__ float4 s_nrm[64];
    __shared__ float4 s_flux[64];
    __shared__ uint32_t s_start[64];
    __shared__ uint32_t s_pre[65];
    const int64_t n = a.hp_count_src ? (int64_t)a.hp_count_src->hitpoints : a.n_hit;
    const int lane = threadIdx.x;
    const int64_t i = (int64_t)blockIdx.x * 64 + lane;
    if ((int64_t)blockIdx.x * 64 >= n) return;  // whole workgroup idle (uniform)
    const bool valid = i < n;
    float4 hp = make_float4(0.0f, 0.0f, 0.0f, 0.0f), hw = hp;
    if (valid) {
        hp = a.hp_pos[i];
        hw = a.hp_w[i];
    }
    const v3f x = xyz(hp);
    const v3f nx = valid ? oct_decode(__float_as_uint(hp.w)) : mk(0.0f, 0.0f, 0.0f);
    uint32_t mb[8];
    const int ns = valid ? own_buckets(a, hp, mb) : 0;
    unsigned long long nb = (unsigned long long)ns, nc = 0, ni = 0;
    for (int k = 0; k < 8; k++)
        if (k < ns) nc += a.off[mb[k] + 1] - a.off[mb[k]];
    uint32_t pending = (1u << ns) - 1u;
    const float r2 = (float)(a.radius * a.radius);
    const float inv_kr = (float)(1.0 / (1.1 * a.radius));
    float ax = 0.0f, ay = 0.0f, az = 0.0f;
    if (lane == 0) s_pre[0] = 0;
    while (true) {
        // ---- collect up to 64 buckets of the wave's union, one per lane ----
        uint32_t my_bucket = 0;
        int nu = 0;
        while (nu < 64) {
            const uint64_t any = __ballot(pending != 0u);
            if (any == 0) break;
            const int leader = __ffsll((unsigned long long)any) - 1;
            uint32_t first = 0;
            {
                const int k0 = pending ? __ffs(pending) - 1 : 0;
                for (int k = 0; k < 8; k++)
                    if (k == k0) first = mb[k];
            }
            const uint32_t b = __shfl(first, leader);
            for (int k = 0; k < 8; k++)
                if (k < ns && mb[k] == b) pending &= ~(1u << k);
            if (lane == nu) my_bucket = b;
            nu++;
        }
        if (nu == 0) break;
        // ---- bucket ranges and their prefix over the lanes ----
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
        // ---- stream the union's photons through LDS ----
        int u0 = 0;
        for (uint32_t base = 0; base < total; base += 64) {
            const uint32_t g = base + (uint32_t)lane;
            if (g < total) {
                int u = u0;
                while (g >= s_pre[u + 1]) u++;
                const uint32_t j = s_start[u] + (g - s_pre[u]);
                const float4 pp = a.ph_pos[j];
                const float4 ff = a.ph_flux[j];
                const v3f d = oct_direction(__float_as_uint(pp.w));
                s_pos[lane] = make_float4(pp.x, pp.y, pp.z, 0.0f);
                s_nrm[lane] = make_float4(d.x, d.y, d.z, 0.0f);
                s_flux[lane] = ff;
            }
            __syncthreads();
            const int m = (total - base) < 64u ? (int)(total - base) : 64;
            for (int k = 0; k < m; k++) {
                const float4 pp = s_pos[k];
                const float dx = x.x - pp.x, dy = x.y - pp.y, dz = x.z - pp.z;
                const float d2 = dx * dx + dy * dy + dz * dz;
                if (d2 <= r2) {
                    const float4 pn = s_nrm[k];
                    if (pn.x * nx.x + pn.y * nx.y + pn.z * nx.z > 0.0f) {
                        const float4 f = s_flux[k];
                        const float wt = 1.0f - sqrtf(d2) * inv_kr;
                        ax += wt * f.x;
                        ay += wt * f.y;
                        az += wt * f.z;
                        ni++;
                    }
                }
            }
            __syncthreads();
            const uint32_t next = base + 64;
            while (u0 < nu - 1 && next >= s_pre[u0 + 1]) u0++;
        }
    }
    if (valid) write_result(a, i, (int64_t)__float_as_uint(hw.w), xyz(hw), ax, ay, az);
    nb = wave_sum(valid ? nb : 0);
    nc = wave_sum(valid ? nc : 0);
    ni = wave_sum(valid ? ni : 0);
    if (lane == 0 && a.counters) {
        atomicAdd(&a.counters->gather_buckets, nb);
        atomicAdd(&a.counters->gather_candidates, nc);
        atomicAdd(&a.counters->gather_inradius, ni);
    }
}

__device__ inline uint32_t spread10(uint32_t v) {
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
    const double inv = 1.0 / radius;
    const double c[3] = {((double)p.x - g.bmin[0]) * inv, ((double)p.y - g.bmin[1]) * inv,
                         ((double)p.z - g.bmin[2]) * inv};
    uint32_t q[3];
    for (int d = 0; d < 3; d++) q[d] = (uint32_t)fmin(fmax(c[d], 0.0), 1023.0);
    keys[i] = (spread10(q[0]) << 2) | (spread10(q[1]) << 1) | spread10(q[2]);
    vals[i] = (uint32_t)i;
}

}  // namespace

static inline unsigned grid1(int64_t n, int b) { return (unsigned)((n + b - 1) / b); }

void launch_gather(hipStream_t s, const GatherArgs& a) {
    // with a device-side count the grid covers the worst case (one per pixel)
    if (a.n_hit <= 0) return;
    if (a.mode == GATHER_LANES)
        hipLaunchKernelGGL(k_gather_lanes, dim3(grid1(a.n_hit, 256)), dim3(256), 0, s, a);
    else
        hipLaunchKernelGGL(k_gather_tiled, dim3(grid1(a.n_hit, 64)), dim3(64), 0, s, a);
}

void launch_hit_keys(hipStream_t s, const float4* hp_pos, const float4* hp_w, int64_t cap, GridParams g,
                     double radius, uint32_t* keys, uint32_t* vals) {
    if (cap <= 0) return;
    hipLaunchKernelGGL(k_hit_keys, dim3(grid1(cap, 256)), dim3(256), 0, s, hp_pos, hp_w, cap, g, radius, keys,
                       vals);
}
