// dipole_cut.hip -- the octree cut of the dipole sum, an option beside the
// exact all-pairs sum of subsurface.hip (pppm_subsurface_cut).
//
// SubsurfaceIntegrator::Octree (sources/subsurface_integrator.cc:69-167):
// over the surface points, each node's box is the tight box of its points and
// its children split that box at the centre, child index
// (x >= mid ? 4 : 0) + (y >= mid ? 2 : 0) + (z >= mid ? 1 : 0).  A node
// stands for its points with
//   position   = luminance-weighted mean of its children's positions,
//   area       = sum of its children's areas,
//   irradiance = mean of its children's irradiances        (:113-142)
// and the sum at x walks the tree, taking a node whole when it is a leaf or
// when area / d^2 < max_error with x outside its box (:147-167).  Points at
// one position share one leaf (the reference would split them forever).
//
// Layout: the tree topology depends only on the points, so it is built once
// on the host (FP64 boxes of the FP32 points, exactly representable) and
// stored in preorder with an escape index per node: the walk needs no stack
// (take the node: jump to its escape; open it: step to the next node), and a
// node's children are i + 1, escape(i + 1), ... below escape(i).  The node
// values depend on the pass's irradiance: one launch per tree level, deepest
// first, recomputes them.  Per node 2 x 16 bytes are read on every visit
// (position + area, irradiance + escape) and the box (2 x 16 bytes) only when
// the error bound would take it.
#include "ctx.h"

#include <functional>

namespace {

constexpr uint32_t kLeafBit = 0x80000000u;

__device__ __forceinline__ float lum(float r, float g, float b) { return 0.2126f * r + 0.7152f * g + 0.0722f * b; }

// node values of one tree level from the irradiance E [point][3] (leaves) or
// from their children (interior nodes, already updated by deeper levels)
__global__ __launch_bounds__(256) void k_cut_update(CutDev t, const int32_t* __restrict__ nodes, int64_t count,
                                                     const float* __restrict__ E) {
    const int64_t k = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (k >= count) return;
    const int32_t i = nodes[k];
    const float4 lo = t.lo[i], hi = t.hi[i];
    const uint32_t esc = __float_as_uint(t.b[i].w);
    const uint32_t n_leaf = __float_as_uint(hi.w);
    float px = 0.0f, py = 0.0f, pz = 0.0f, ir = 0.0f, ig = 0.0f, ib = 0.0f;
    if (esc & kLeafBit) {
        const uint32_t first = __float_as_uint(lo.w);
        for (uint32_t j = 0; j < n_leaf; j++) {
            const uint32_t p = t.perm[first + j];
            ir += E[3 * (int64_t)p];
            ig += E[3 * (int64_t)p + 1];
            ib += E[3 * (int64_t)p + 2];
        }
        const float inv = 1.0f / (float)n_leaf;
        ir *= inv;
        ig *= inv;
        ib *= inv;
        px = lo.x;
        py = lo.y;
        pz = lo.z;
    } else {
        const int32_t end = (int32_t)(esc & ~kLeafBit);
        float w = 0.0f;
        int nc = 0;
        for (int32_t j = i + 1; j < end; j = (int32_t)(__float_as_uint(t.b[j].w) & ~kLeafBit)) {
            const float4 ca = t.a[j], cb = t.b[j];
            const float lw = lum(cb.x, cb.y, cb.z);
            px += lw * ca.x;
            py += lw * ca.y;
            pz += lw * ca.z;
            ir += cb.x;
            ig += cb.y;
            ib += cb.z;
            w += lw;
            nc++;
        }
        if (w > 0.0f) {
            px /= w;
            py /= w;
            pz /= w;
        }
        const float inv = 1.0f / (float)nc;
        ir *= inv;
        ig *= inv;
        ib *= inv;
    }
    t.a[i] = make_float4(px, py, pz, t.a[i].w);
    t.b[i] = make_float4(ir, ig, ib, __uint_as_float(esc));
}

// R_d per channel at squared distance s from a source at depth z (see
// subsurface.hip): z (sigma_tr d + 1) e^{-sigma_tr d} / d^3
__device__ __forceinline__ float cut_source(float s, float trl, float zt, float z) {
    const float id = __builtin_amdgcn_rsqf(s);
    const float d = s * id;
    const float e = __builtin_amdgcn_exp2f(-trl * d);
    return __builtin_fmaf(zt, d, z) * e * (id * id * id);
}

// one lane per eye path that entered a translucent surface: the stackless
// walk of the cut, FP32 terms with a Kahan-compensated sum per channel
__global__ __launch_bounds__(256) void k_dipole_cut(DipoleArgs a, CutDev t) {
    const int64_t n = a.count_src ? (int64_t)a.count_src->sss_hits : a.n_hit;
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const float4 hp = a.hit_pos[i];
    const DipoleDev D = a.dip[__float_as_int(hp.w)];
    const float log2e = 1.4426950408889634f;
    float zp2[3], zn2[3], trl[3], ztp[3], ztn[3];
#pragma unroll
    for (int c = 0; c < 3; c++) {
        zp2[c] = D.zp[c] * D.zp[c];
        zn2[c] = D.zn[c] * D.zn[c];
        trl[c] = D.tr[c] * log2e;
        ztp[c] = D.zp[c] * D.tr[c];
        ztn[c] = D.zn[c] * D.tr[c];
    }
    float s[3] = {0.0f, 0.0f, 0.0f}, comp[3] = {0.0f, 0.0f, 0.0f};
    int32_t k = 0;
    while (k < t.n_nodes) {
        const float4 na = t.a[k], nb = t.b[k];
        const uint32_t esc = __float_as_uint(nb.w);
        const float dx = hp.x - na.x, dy = hp.y - na.y, dz = hp.z - na.z;
        const float d2 = __builtin_fmaf(dx, dx, __builtin_fmaf(dy, dy, dz * dz));
        bool take = (esc & kLeafBit) != 0;
        if (!take && na.w < t.max_error * d2) {
            const float4 lo = t.lo[k], hi = t.hi[k];
            const bool inside = lo.x <= hp.x && hp.x <= hi.x && lo.y <= hp.y && hp.y <= hi.y && lo.z <= hp.z &&
                                hp.z <= hi.z;
            take = !inside;
        }
        if (!take) {
            k++;
            continue;
        }
        const float ir[3] = {nb.x, nb.y, nb.z};
#pragma unroll
        for (int c = 0; c < 3; c++) {
            const float r = cut_source(d2 + zp2[c], trl[c], ztp[c], D.zp[c]) -
                            cut_source(d2 + zn2[c], trl[c], ztn[c], D.zn[c]);
            const float term = fmaxf(D.k[c] * r, 0.0f) * ir[c] * na.w;
            const float y = term - comp[c];
            const float u = s[c] + y;
            comp[c] = (u - s[c]) - y;
            s[c] = u;
        }
        k = (int32_t)(esc & ~kLeafBit);
    }
    const float4 w = a.hit_w[i];
    const float Lo[3] = {w.x * D.out * s[0], w.y * D.out * s[1], w.z * D.out * s[2]};
    const int64_t pix = (int64_t)__float_as_uint(w.w);
    if (a.img_sum) {
        for (int c = 0; c < 3; c++) kahan_add(a.img_sum, a.img_comp, 3 * pix + c, Lo[c]);
    } else {
        for (int c = 0; c < 3; c++) a.out_rgb[3 * i + c] = Lo[c];
    }
}

}  // namespace

void launch_cut_update(hipStream_t s, const CutDev& t, const int32_t* nodes, int64_t count, const float* E) {
    if (count <= 0) return;
    hipLaunchKernelGGL(k_cut_update, dim3((unsigned)((count + 255) / 256)), dim3(256), 0, s, t, nodes, count, E);
}

void launch_dipole_cut(hipStream_t s, const DipoleArgs& a, const CutDev& t) {
    if (a.n_hit <= 0 || t.n_nodes <= 0) return;
    hipLaunchKernelGGL(k_dipole_cut, dim3((unsigned)((a.n_hit + 255) / 256)), dim3(256), 0, s, a, t);
}

namespace pppmi {

// The tree over n points (xyz of pts[k], stride 4 floats) in preorder.
void cut_build(const float4* pts, int64_t n, double area, CutHost* out) {
    out->lo.clear();
    out->hi.clear();
    out->a.clear();
    out->esc.clear();
    out->perm.clear();
    out->depth.clear();
    if (n <= 0) return;
    std::vector<uint32_t> idx((size_t)n), tmp((size_t)n);
    for (int64_t k = 0; k < n; k++) idx[(size_t)k] = (uint32_t)k;
    out->perm.reserve((size_t)n);
    std::function<void(size_t, size_t, int32_t)> rec = [&](size_t b, size_t m, int32_t depth) {
        double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
        for (size_t k = b; k < b + m; k++) {
            const float4 p = pts[idx[k]];
            const double v[3] = {p.x, p.y, p.z};
            for (int d = 0; d < 3; d++) {
                lo[d] = std::min(lo[d], v[d]);
                hi[d] = std::max(hi[d], v[d]);
            }
        }
        const size_t id = out->lo.size();
        out->lo.push_back(make_float4((float)lo[0], (float)lo[1], (float)lo[2], 0.0f));
        out->hi.push_back(make_float4((float)hi[0], (float)hi[1], (float)hi[2], 0.0f));
        out->a.push_back(make_float4(0.0f, 0.0f, 0.0f, (float)(area * (double)m)));
        out->esc.push_back(0u);
        out->depth.push_back(depth);
        if (m == 1 || (lo[0] == hi[0] && lo[1] == hi[1] && lo[2] == hi[2])) {
            const uint32_t first = (uint32_t)out->perm.size(), cnt = (uint32_t)m;
            for (size_t k = b; k < b + m; k++) out->perm.push_back(idx[k]);
            std::memcpy(&out->lo[id].w, &first, 4);
            std::memcpy(&out->hi[id].w, &cnt, 4);
            out->esc[id] = (uint32_t)out->lo.size() | kLeafBit;
            return;
        }
        double mid[3];
        for (int d = 0; d < 3; d++) mid[d] = (lo[d] + hi[d]) * 0.5;
        auto child = [&](uint32_t q) {
            const float4 p = pts[q];
            return ((double)p.x < mid[0] ? 0 : 4) + ((double)p.y < mid[1] ? 0 : 2) + ((double)p.z < mid[2] ? 0 : 1);
        };
        size_t cnt[8] = {0}, start[8], fill[8];
        for (size_t k = b; k < b + m; k++) cnt[child(idx[k])]++;
        start[0] = b;
        for (int c = 1; c < 8; c++) start[c] = start[c - 1] + cnt[c - 1];
        std::memcpy(fill, start, sizeof fill);
        for (size_t k = b; k < b + m; k++) tmp[fill[child(idx[k])]++] = idx[k];
        std::memcpy(&idx[b], &tmp[b], sizeof(uint32_t) * m);
        for (int c = 0; c < 8; c++)
            if (cnt[c]) rec(start[c], cnt[c], depth + 1);
        out->esc[id] = (uint32_t)out->lo.size();
    };
    rec(0, (size_t)n, 0);
}

hipError_t cut_upload(CutBufs* t, const CutHost& h) {
    const size_t nn = h.lo.size();
    int32_t max_depth = 0;
    for (int32_t d : h.depth) max_depth = std::max(max_depth, d);
    // nodes grouped by depth, each level in preorder
    t->level_off.assign((size_t)max_depth + 2, 0);
    for (int32_t d : h.depth) t->level_off[(size_t)d + 1]++;
    for (size_t d = 1; d < t->level_off.size(); d++) t->level_off[d] += t->level_off[d - 1];
    std::vector<int32_t> lv(nn);
    std::vector<int64_t> fill(t->level_off.begin(), t->level_off.end() - 1);
    for (size_t i = 0; i < nn; i++) lv[(size_t)fill[(size_t)h.depth[i]]++] = (int32_t)i;
    std::vector<float4> b(nn);
    for (size_t i = 0; i < nn; i++) {
        float e;
        std::memcpy(&e, &h.esc[i], 4);
        b[i] = make_float4(0.0f, 0.0f, 0.0f, e);
    }
    t->n_nodes = 0;
    hipError_t e;
    if ((e = t->a.ensure(sizeof(float4) * nn)) != hipSuccess) return e;
    if ((e = t->b.ensure(sizeof(float4) * nn)) != hipSuccess) return e;
    if ((e = t->lo.ensure(sizeof(float4) * nn)) != hipSuccess) return e;
    if ((e = t->hi.ensure(sizeof(float4) * nn)) != hipSuccess) return e;
    if ((e = t->perm.ensure(sizeof(uint32_t) * h.perm.size())) != hipSuccess) return e;
    if ((e = t->levels.ensure(sizeof(int32_t) * nn)) != hipSuccess) return e;
    if ((e = hipMemcpy(t->a.p, h.a.data(), sizeof(float4) * nn, hipMemcpyHostToDevice)) != hipSuccess) return e;
    if ((e = hipMemcpy(t->b.p, b.data(), sizeof(float4) * nn, hipMemcpyHostToDevice)) != hipSuccess) return e;
    if ((e = hipMemcpy(t->lo.p, h.lo.data(), sizeof(float4) * nn, hipMemcpyHostToDevice)) != hipSuccess) return e;
    if ((e = hipMemcpy(t->hi.p, h.hi.data(), sizeof(float4) * nn, hipMemcpyHostToDevice)) != hipSuccess) return e;
    if ((e = hipMemcpy(t->perm.p, h.perm.data(), sizeof(uint32_t) * h.perm.size(), hipMemcpyHostToDevice)) !=
        hipSuccess)
        return e;
    if ((e = hipMemcpy(t->levels.p, lv.data(), sizeof(int32_t) * nn, hipMemcpyHostToDevice)) != hipSuccess) return e;
    t->n_nodes = (int32_t)nn;
    return hipSuccess;
}

CutDev cut_dev(const CutBufs& t, double max_error) {
    CutDev d;
    d.a = t.a.as<float4>();
    d.b = t.b.as<float4>();
    d.lo = t.lo.as<float4>();
    d.hi = t.hi.as<float4>();
    d.perm = t.perm.as<uint32_t>();
    d.n_nodes = t.n_nodes;
    d.max_error = (float)max_error;
    return d;
}

// node values from this pass's irradiance, deepest level first
void cut_refresh(const CutBufs& t, hipStream_t s, const float* E, double max_error) {
    if (t.n_nodes <= 0) return;
    const CutDev d = cut_dev(t, max_error);
    const int32_t* lv = t.levels.as<int32_t>();
    for (size_t k = t.level_off.size() - 1; k-- > 0;)
        launch_cut_update(s, d, lv + t.level_off[k], t.level_off[k + 1] - t.level_off[k], E);
}

}  // namespace pppmi
