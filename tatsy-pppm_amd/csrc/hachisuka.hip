// hachisuka.hip -- progressive photon mapping with per-point radii (Hachisuka,
// Ogaki and Jensen 2008) on the GPU hash grid: the render method of
// ProgressivePhotonMapping (sources/progressive_photon_mapping.cc:22-287),
// selected by pppm_render_params.method = PPPM_METHOD_HACHISUKA.
//
// Each pixel keeps a render point across passes: its radius^2, its photon
// count and its accumulated flux.  A pass traces one eye path per pixel to its
// first Lambertian surface (no Russian roulette, as :300-343), then photons,
// then updates every render point with the photons of this pass within its
// radius, one accepted photon at a time:
//     g = (n alpha + alpha) / (n alpha + 1),  r^2 <- r^2 g,  n <- n + 1,
//     tau <- (tau + W Phi) g                                          (:244-258)
// (with N = n alpha accepted photons this is the standard update
// (N + alpha) / (N + 1)).  The reference lets every photon update the points
// of its hash bucket under a lock, so the order of updates depends on its
// threads.  Here a point walks its candidates itself: the deduplicated
// buckets of the cells its box [x - R, x + R] meets (R the pass's initial
// radius), in z, y, x order, each bucket in photon order -- one fixed order,
// no atomics, the oracle's order.  The test in FP64, like the update.
// The image is (emission + tau / (pi r^2)) / passes (:83-86).
#include "ctx.h"

namespace {

// per-block min / max of the hit points' positions
__global__ __launch_bounds__(256) void k_ppm_bounds(const float4* __restrict__ hp_pos, const float4* __restrict__ hp_w,
                                                    int64_t np, float* __restrict__ part) {
    __shared__ float s[6][256];
    float lo[3] = {3.0e38f, 3.0e38f, 3.0e38f}, hi[3] = {-3.0e38f, -3.0e38f, -3.0e38f};
    for (int64_t p = (int64_t)blockIdx.x * 256 + threadIdx.x; p < np; p += (int64_t)gridDim.x * 256) {
        if (__float_as_uint(hp_w[p].w) == NO_HITPOINT) continue;
        const float4 x = hp_pos[p];
        const float xs[3] = {x.x, x.y, x.z};
        for (int d = 0; d < 3; d++) {
            lo[d] = fminf(lo[d], xs[d]);
            hi[d] = fmaxf(hi[d], xs[d]);
        }
    }
    for (int d = 0; d < 3; d++) {
        s[d][threadIdx.x] = lo[d];
        s[3 + d][threadIdx.x] = hi[d];
    }
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off)
            for (int d = 0; d < 3; d++) {
                s[d][threadIdx.x] = fminf(s[d][threadIdx.x], s[d][threadIdx.x + off]);
                s[3 + d][threadIdx.x] = fmaxf(s[3 + d][threadIdx.x], s[3 + d][threadIdx.x + off]);
            }
        __syncthreads();
    }
    if (threadIdx.x < 6) part[6 * blockIdx.x + threadIdx.x] = s[threadIdx.x][0];
}

struct PpmArgs {
    const float4* hp_pos;  // per pixel: xyz, oriented normal code
    const float4* hp_w;    // weight (throughput rho / pi) rgb, pixel id (NO_HITPOINT: none)
    int64_t np;
    const float4* ph_a;    // photons in bucket order: xyz, normal code
    const float4* ph_b;    // flux rgb, key
    const uint32_t* off;   // bucket offsets [S + 1]
    GridParams g;
    double irad;           // initial radius of this pass: the half-width of a point's box
    double alpha;
    double* r2;            // per pixel
    uint32_t* n;           // accepted photons per pixel
    double* flux;          // tau per pixel [np][3]
    Counters* counters;
};

__global__ __launch_bounds__(256) void k_ppm_update(PpmArgs a) {
    unsigned long long cand = 0, acc = 0, nb = 0;
    for (int64_t p = (int64_t)blockIdx.x * 256 + threadIdx.x; p < a.np; p += (int64_t)gridDim.x * 256) {
        const float4 w = a.hp_w[p];
        if (__float_as_uint(w.w) == NO_HITPOINT) continue;
        const float4 x = a.hp_pos[p];
        const v3f nx = oct_decode(__float_as_uint(x.w));
        double r2 = a.r2[p], f0 = a.flux[3 * p], f1 = a.flux[3 * p + 1], f2 = a.flux[3 * p + 2];
        uint32_t cnt = a.n[p];
        if (cnt == 0) {  // no photon yet: the pass's initial radius (:125-131)
            r2 = a.irad * a.irad;
            f0 = f1 = f2 = 0.0;
        }
        const double xs[3] = {x.x, x.y, x.z};
        int32_t lo[3], hi[3];
        for (int d = 0; d < 3; d++) {
            lo[d] = grid_cell(xs[d] - a.irad, a.g.bmin[d], a.g.scale);
            hi[d] = min(grid_cell(xs[d] + a.irad, a.g.bmin[d], a.g.scale), lo[d] + 1);
        }
        uint32_t seen[8];
        int ns = 0;
        for (int32_t iz = lo[2]; iz <= hi[2]; iz++)
            for (int32_t iy = lo[1]; iy <= hi[1]; iy++)
                for (int32_t ix = lo[0]; ix <= hi[0]; ix++) {
                    const uint32_t h = grid_hash(ix, iy, iz, a.g.size);
                    bool dup = false;
                    for (int k = 0; k < ns; k++) dup |= seen[k] == h;
                    if (dup) continue;
                    seen[ns++] = h;
                    nb++;
                    const uint32_t e = a.off[h + 1];
                    for (uint32_t j = a.off[h]; j < e; j++) {
                        const float4 pa = a.ph_a[j];
                        cand++;
                        const v3f np_ = oct_decode(__float_as_uint(pa.w));
                        if (!(dot(np_, nx) > 0.0f)) continue;
                        double d2;
                        {
#pragma clang fp contract(off)
                            const double dx = xs[0] - (double)pa.x, dy = xs[1] - (double)pa.y, dz = xs[2] - (double)pa.z;
                            d2 = dx * dx + dy * dy + dz * dz;
                        }
                        if (!(d2 <= r2)) continue;
                        const float4 pb = a.ph_b[j];
                        const double na = (double)cnt * a.alpha;
                        const double g = (na + a.alpha) / (na + 1.0);
                        r2 *= g;
                        cnt++;
                        f0 = (f0 + (double)w.x * (double)pb.x) * g;
                        f1 = (f1 + (double)w.y * (double)pb.y) * g;
                        f2 = (f2 + (double)w.z * (double)pb.z) * g;
                        acc++;
                    }
                }
        a.r2[p] = r2;
        a.n[p] = cnt;
        a.flux[3 * p] = f0;
        a.flux[3 * p + 1] = f1;
        a.flux[3 * p + 2] = f2;
    }
    cand = wave_sum(cand);
    acc = wave_sum(acc);
    nb = wave_sum(nb);
    if ((threadIdx.x & 63) == 0) {
        atomicAdd(&a.counters->gather_candidates, cand);
        atomicAdd(&a.counters->gather_inradius, acc);
        atomicAdd(&a.counters->gather_buckets, nb);
    }
}

// image = ((dsum - dcomp) + tau / (pi r^2)) / passes, FP64
__global__ void k_ppm_image(const float* __restrict__ dsum, const float* __restrict__ dcomp,
                            const double* __restrict__ r2, const uint32_t* __restrict__ n,
                            const double* __restrict__ flux, int64_t np, double inv_passes, double* out) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= np) return;
    for (int k = 0; k < 3; k++) {
        const double e = (double)dsum[3 * p + k] - (double)dcomp[3 * p + k];
        const double t = n[p] > 0 ? flux[3 * p + k] / (3.14159265358979323846 * r2[p]) : 0.0;
        out[3 * p + k] = (e + t) * inv_passes;
    }
}

}  // namespace

namespace pppmi {

hipError_t hachisuka_prepare(Ctx* c) {
    const size_t np = (size_t)c->cam.width * c->cam.height;
    hipError_t e;
    if ((e = c->ppm_r2.ensure(sizeof(double) * np)) != hipSuccess) return e;
    if ((e = c->ppm_n.ensure(sizeof(uint32_t) * np)) != hipSuccess) return e;
    if ((e = c->ppm_flux.ensure(sizeof(double) * 3 * np)) != hipSuccess) return e;
    if ((e = c->ppm_bounds.ensure(sizeof(float) * 6 * 1024)) != hipSuccess) return e;
    launch_zero_u32(c->stream, c->ppm_n.p, (int64_t)np);
    launch_zero_u32(c->stream, c->ppm_r2.p, 2 * (int64_t)np);
    launch_zero_u32(c->stream, c->ppm_flux.p, 6 * (int64_t)np);
    return hipGetLastError();
}

// One pass of the Hachisuka method.  Returns an error message (empty: fine).
std::string hachisuka_pass(Ctx* c, pppm_pass_stats* st) {
    hipStream_t s = c->stream;
    const int32_t pass = c->passes_done + 1;
    const int64_t np = (int64_t)c->cam.width * c->cam.height;
    Counters* dc = c->counters.as<Counters>();
    launch_zero_u32(s, dc, sizeof(Counters) / 4);
    // eye paths to their first Lambertian surface, no roulette; what they see
    // of emitters and the environment is Kahan-added to the direct image
    EyePassArgs ea = eye_args(c, (uint32_t)pass);
    ea.roulette = 0;
    launch_eye_pass(s, ea, c->cu > 0 ? c->cu * EYE_BLOCKS_PER_CU : 2048);
    // the initial radius of this pass from the hit points' extent (:116-125)
    const int blocks = 1024;
    hipLaunchKernelGGL(k_ppm_bounds, dim3(blocks), dim3(256), 0, s, c->hp_pos.as<float4>(), c->hp_w.as<float4>(), np,
                       c->ppm_bounds.as<float>());
    std::vector<float> part(6 * blocks);
    if (hipMemcpyAsync(part.data(), c->ppm_bounds.p, sizeof(float) * part.size(), hipMemcpyDeviceToHost, s) !=
        hipSuccess)
        return "bounds copy failed";
    // photons on the same stream, then their count
    PhotonPassArgs pa;
    pa.sc = c->scene_dev();
    pa.n_photons = c->rp.photons;
    pa.cap = c->cap;
    pa.seed = c->rp.seed;
    pa.pass = (uint32_t)pass;
    pa.scene_id = c->rp.scene_id;
    pa.max_bounces = c->rp.max_bounces;
    pa.ray_eps = (float)c->rp.ray_eps;
    pa.ph_pos = c->ph_rec.as<float4>();
    pa.ph_flux = c->ph_rec.as<float4>() + 1;
    pa.rec_stride = 2;
    pa.counters = dc;
    pa.prop.cdf = nullptr;
    pa.prop.lambda = 1.0f;
    pa.prop.side = 0;
    pa.irr_pos = nullptr;
    pa.irr_flux = nullptr;
    pa.irr_cap = 0;
    pa.count_traversal = c->count_traversal;
    pa.rec_rank = c->rec_rank.as<uint8_t>();
    pa.rec_count = c->rec_count.as<uint32_t>();
    launch_photon_pass(s, pa, c->cu > 0 ? c->cu * PHOTON_BLOCKS_PER_CU : 2048);
    Counters hc;
    if (hipMemcpyAsync(&hc, dc, sizeof(Counters), hipMemcpyDeviceToHost, s) != hipSuccess) return "count copy failed";
    if (hipStreamSynchronize(s) != hipSuccess) return std::string("pass: ") + hipGetErrorString(hipGetLastError());
    float lo[3] = {3.0e38f, 3.0e38f, 3.0e38f}, hi[3] = {-3.0e38f, -3.0e38f, -3.0e38f};
    for (int b = 0; b < blocks; b++)
        for (int d = 0; d < 3; d++) {
            lo[d] = std::min(lo[d], part[6 * b + d]);
            hi[d] = std::max(hi[d], part[6 * b + 3 + d]);
        }
    const int64_t n = (int64_t)hc.photons_stored;
    if (st) {
        std::memset(st, 0, sizeof(*st));
        st->pass_index = pass;
        st->photons_emitted = c->rp.photons;
        st->photons_stored = n;
        st->hitpoints = (int64_t)hc.hitpoints;
        st->photon_segments = (int64_t)hc.photon_segments;
        st->eye_segments = (int64_t)hc.eye_segments;
    }
    if (n > c->cap) return "photon buffer holds " + std::to_string(c->cap) + " records, the pass stored " + std::to_string(n);
    if (hc.hitpoints == 0) return "";  // no render point this pass: nothing to update
    // irad = mean extent / mean image side * 8 (:120-121)
    const double ext = (((double)hi[0] - (double)lo[0]) + ((double)hi[1] - (double)lo[1]) +
                        ((double)hi[2] - (double)lo[2])) / 3.0;
    double irad = ext / ((c->cam.width + c->cam.height) / 2.0) * 8.0;
    if (!(irad > 0.0)) irad = c->rp.radius1;  // a single hit point has no extent
    if (st) st->radius = irad;
    // photons in (photon index, bounce) order, then a hash grid of cells 2 irad
    uint32_t S = c->rp.hash_size > 0 ? (uint32_t)c->rp.hash_size : (uint32_t)(n > 0 ? n : 1);
    GridParams g;
    for (int d = 0; d < 3; d++) g.bmin[d] = c->bmin[d] - irad;
    g.scale = 1.0 / (2.0 * irad);
    g.size = S;
    g.sub_bits = 0;
    if (c->off.ensure(sizeof(uint32_t) * ((size_t)S + 1)) != hipSuccess) return "offset buffer allocation failed";
    uint32_t* perm1 = origin_order(c, s, n, g, nullptr);
    uint32_t* sorted = build_grid(c, nullptr, nullptr, perm1, n, g, 4, nullptr, nullptr, true);
    launch_gather_layout(s, c->ph_rec.as<float4>(), c->ph_rec.as<float4>() + 1, nullptr, sorted, n,
                         c->ph_a.as<float4>(), c->ph_b.as<float4>(), nullptr, nullptr, 2);
    PpmArgs u;
    u.hp_pos = c->hp_pos.as<float4>();
    u.hp_w = c->hp_w.as<float4>();
    u.np = np;
    u.ph_a = c->ph_a.as<float4>();
    u.ph_b = c->ph_b.as<float4>();
    u.off = c->off.as<uint32_t>();
    u.g = g;
    u.irad = irad;
    u.alpha = c->rp.alpha;
    u.r2 = c->ppm_r2.as<double>();
    u.n = c->ppm_n.as<uint32_t>();
    u.flux = c->ppm_flux.as<double>();
    u.counters = dc;
    hipLaunchKernelGGL(k_ppm_update, dim3((unsigned)std::min<int64_t>((np + 255) / 256, c->cu > 0 ? 16 * c->cu : 4096)),
                       dim3(256), 0, s, u);
    if (hipMemcpyAsync(&hc, dc, sizeof(Counters), hipMemcpyDeviceToHost, s) != hipSuccess) return "count copy failed";
    if (hipStreamSynchronize(s) != hipSuccess) return std::string("update: ") + hipGetErrorString(hipGetLastError());
    if (st) {
        st->hash_size = (int32_t)S;
        st->gather_buckets = (int64_t)hc.gather_buckets;
        st->gather_candidates = (int64_t)hc.gather_candidates;
        st->gather_inradius = (int64_t)hc.gather_inradius;
    }
    return "";
}

hipError_t hachisuka_image(Ctx* c, double* out_dev) {
    const int64_t np = (int64_t)c->cam.width * c->cam.height;
    const double inv = c->passes_done > 0 ? 1.0 / (double)c->passes_done : 0.0;
    hipLaunchKernelGGL(k_ppm_image, dim3((unsigned)((np + 255) / 256)), dim3(256), 0, c->stream, c->img_dsum.as<float>(),
                       c->img_dcomp.as<float>(), c->ppm_r2.as<double>(), c->ppm_n.as<uint32_t>(),
                       c->ppm_flux.as<double>(), np, inv, out_dev);
    return hipGetLastError();
}

}  // namespace pppmi
