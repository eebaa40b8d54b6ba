// sss_host.hip -- subsurface scattering, host side: surface points, the
// per-pass irradiance of the points, and the pppm_subsurface_* / dipole entry
// points (SubsurfaceIntegrator, sources/subsurface_integrator.cc:180-365;
// DipoleBSSRDF, sources/bssrdf.cc:33-104).
#include "ctx.h"

using namespace pppmi;

namespace pppmi {

// ---------------------------------------------------------------------------
// Subsurface scattering: host side
// ---------------------------------------------------------------------------
// F_dr, the rational fit of sources/bssrdf.cc:33-42
double dipole_fdr(double eta) {
    const double e2 = eta * eta, e3 = e2 * eta;
    if (eta >= 1.0) return (-1.439 + 0.7099 * eta + 0.6681 * e2 + 0.0636 * e3) / e2;
    return (-0.4399 * e3 + 0.7099 * e2 - 0.3319 * eta + 0.0636) / e3;
}

uint32_t morton_bits(uint32_t v) {  // 10 bits spread 3 apart
    v &= 0x3FFu;
    v = (v | (v << 16)) & 0x030000FFu;
    v = (v | (v << 8)) & 0x0300F00Fu;
    v = (v | (v << 4)) & 0x030C30C3u;
    v = (v | (v << 2)) & 0x09249249u;
    return v;
}

// Surface points on the BSSRDF triangles (the reference draws a Poisson disk
// every pass, sources/sampler.cc:36-101, with dA = pi (r / 2)^2): n points
// placed once by systematic sampling of the area CDF, uniform barycentrics
// inside each triangle, so every point stands for area A / n.  The default
// count gives the reference's dA (r = sqrt(sum of BSSRDF areas / all
// triangles), ppm_probabilistic.cc:37-46), at most 2^20.  Points are stored in
// Morton order.
uint64_t splitmix64(uint64_t* st);

hipError_t subsurface_setup(Ctx* c, const float* tri_h /* may be null: read back */) {
    c->sss_present = false;
    c->sss_n = 0;
    if (c->n_dips == 0) return hipSuccess;
    std::vector<float> tri((size_t)c->n_tris * 9);
    std::vector<int32_t> mat((size_t)c->n_tris);
    hipError_t e;
    if (tri_h) std::memcpy(tri.data(), tri_h, sizeof(float) * tri.size());
    else if ((e = hipMemcpy(tri.data(), c->tri9.p, sizeof(float) * tri.size(), hipMemcpyDeviceToHost)) != hipSuccess)
        return e;
    if ((e = hipMemcpy(mat.data(), c->tri_mat.p, sizeof(int32_t) * mat.size(), hipMemcpyDeviceToHost)) != hipSuccess)
        return e;
    std::vector<int64_t> ids;
    std::vector<double> cdf;
    double total = 0.0;
    for (int64_t t = 0; t < c->n_tris; t++) {
        if (c->h_sss_index[(size_t)mat[(size_t)t]] < 0) continue;
        const float* v = &tri[(size_t)t * 9];
        const double e1[3] = {v[3] - v[0], v[4] - v[1], v[5] - v[2]};
        const double e2[3] = {v[6] - v[0], v[7] - v[1], v[8] - v[2]};
        double cr[3];
        cross3(e1, e2, cr);
        const double a = 0.5 * std::sqrt(cr[0] * cr[0] + cr[1] * cr[1] + cr[2] * cr[2]);
        if (!(a > 0.0)) continue;
        total += a;
        ids.push_back(t);
        cdf.push_back(total);
    }
    if (ids.empty()) return hipSuccess;
    int64_t n = c->sss_points_req;
    if (n <= 0) {
        const double r = std::sqrt(total / (double)c->n_tris);
        const double dA = M_PI * 0.25 * r * r;
        n = std::min<int64_t>((int64_t)std::ceil(total / dA), 1LL << 20);
    }
    std::vector<float4> pts((size_t)n);
    uint64_t st = 0x55D1CE5EEDull;
    size_t cur = 0;
    for (int64_t k = 0; k < n; k++) {
        const double target = ((double)k + 0.5) / (double)n * total;
        while (cur + 1 < cdf.size() && cdf[cur] < target) cur++;
        const float* v = &tri[(size_t)ids[cur] * 9];
        double u1 = (double)(splitmix64(&st) >> 11) * (1.0 / 9007199254740992.0);
        double u2 = (double)(splitmix64(&st) >> 11) * (1.0 / 9007199254740992.0);
        if (u1 + u2 >= 1.0) {
            u1 = 1.0 - u1;
            u2 = 1.0 - u2;
        }
        double p[3];
        for (int d = 0; d < 3; d++) p[d] = v[d] + u1 * (v[3 + d] - v[d]) + u2 * (v[6 + d] - v[d]);
        const double e1[3] = {v[3] - v[0], v[4] - v[1], v[5] - v[2]};
        const double e2[3] = {v[6] - v[0], v[7] - v[1], v[8] - v[2]};
        double nn[3];
        cross3(e1, e2, nn);
        const double ln = std::sqrt(nn[0] * nn[0] + nn[1] * nn[1] + nn[2] * nn[2]);
        const v3f nf = mk((float)(nn[0] / ln), (float)(nn[1] / ln), (float)(nn[2] / ln));
        pts[(size_t)k] = make_float4((float)p[0], (float)p[1], (float)p[2], oct_encode_host(nf));
    }
    // Morton order over the bounding box of the points
    float lo[3] = {1e30f, 1e30f, 1e30f}, hi[3] = {-1e30f, -1e30f, -1e30f};
    for (const float4& q : pts) {
        const float xs[3] = {q.x, q.y, q.z};
        for (int d = 0; d < 3; d++) {
            lo[d] = std::min(lo[d], xs[d]);
            hi[d] = std::max(hi[d], xs[d]);
        }
    }
    std::vector<std::pair<uint32_t, uint32_t>> key((size_t)n);
    for (int64_t k = 0; k < n; k++) {
        const float xs[3] = {pts[(size_t)k].x, pts[(size_t)k].y, pts[(size_t)k].z};
        uint32_t code = 0;
        for (int d = 0; d < 3; d++) {
            const float ext = hi[d] - lo[d];
            const uint32_t q = ext > 0.0f ? (uint32_t)std::min(1023.0f, (xs[d] - lo[d]) / ext * 1024.0f) : 0u;
            code |= morton_bits(q) << d;
        }
        key[(size_t)k] = {code, (uint32_t)k};
    }
    std::stable_sort(key.begin(), key.end(),
                     [](const std::pair<uint32_t, uint32_t>& x, const std::pair<uint32_t, uint32_t>& y) {
                         return x.first < y.first;
                     });
    std::vector<float4> sorted((size_t)n);
    for (int64_t k = 0; k < n; k++) sorted[(size_t)k] = pts[(size_t)key[(size_t)k].second];
    std::vector<float4> wts((size_t)n);
    for (int64_t k = 0; k < n; k++) {
        const uint32_t id = (uint32_t)k;
        float idf;
        std::memcpy(&idf, &id, 4);
        wts[(size_t)k] = make_float4((float)(1.0 / M_PI), (float)(1.0 / M_PI), (float)(1.0 / M_PI), idf);
    }
    if ((e = c->sss_ptw.ensure(sizeof(float4) * (size_t)n)) != hipSuccess) return e;
    if ((e = hipMemcpy(c->sss_ptw.p, wts.data(), sizeof(float4) * (size_t)n, hipMemcpyHostToDevice)) != hipSuccess)
        return e;
    if ((e = c->sss_pts.ensure(sizeof(float4) * (size_t)n)) != hipSuccess) return e;
    if ((e = c->sss_E.ensure(sizeof(float) * 3 * (size_t)n)) != hipSuccess) return e;
    if ((e = hipMemcpy(c->sss_pts.p, sorted.data(), sizeof(float4) * (size_t)n, hipMemcpyHostToDevice)) != hipSuccess)
        return e;
    c->sss_n = n;
    c->sss_area = total / (double)n;
    c->cut.n_nodes = 0;
    if (c->sss_cut > 0.0) {  // the octree over the points, built once per render
        CutHost h;
        cut_build(sorted.data(), n, c->sss_area, &h);
        if ((e = cut_upload(&c->cut, h)) != hipSuccess) return e;
    }
    c->sss_present = true;
    return hipSuccess;
}

// E_j of the surface points from the n_irr irradiance photons of this pass:
// ordered by origin key, a hash grid at the pass radius, the compact gather
// layout (in ph_a / ph_b, rewritten later by the main photons), then the
// lanes gather with weight 1 / pi writing E [point][3]
hipError_t subsurface_irradiance(Ctx* c, int64_t n_irr, double radius) {
    hipStream_t s = c->stream;
    if (n_irr <= 0) {
        launch_zero_u32(s, c->sss_E.p, 3 * c->sss_n);
        return hipGetLastError();
    }
    uint32_t* u0 = c->u[0].as<uint32_t>();
    uint32_t* u1 = c->u[1].as<uint32_t>();
    launch_origin_keys(s, c->irr_flux.as<float4>(), n_irr, u0, u1);
    const int obits = bits_for((uint64_t)(c->rp.photons - 1));
    const bool alt = radix_sort_pairs(s, u0, u1, c->u[2].as<uint32_t>(), c->u[3].as<uint32_t>(),
                                      c->hist.as<uint32_t>(), n_irr, obits);
    const int keep = alt ? 3 : 1;
    const uint32_t* perm = c->u[keep].as<uint32_t>();
    GridParams g;
    for (int d = 0; d < 3; d++) g.bmin[d] = c->grid_bmin[d];
    g.scale = 1.0 / (2.0 * radius);
    g.size = c->rp.hash_size > 0 ? (uint32_t)c->rp.hash_size : (uint32_t)n_irr;
    const uint32_t* sorted = build_grid(c, c->irr_pos.as<float4>(), nullptr, perm, n_irr, g, keep);
    launch_gather_layout(s, c->irr_pos.as<float4>(), c->irr_flux.as<float4>(), perm, sorted, n_irr,
                         c->ph_a.as<float4>(), c->ph_b.as<float4>(), nullptr);
    GatherArgs ga;
    std::memset(&ga, 0, sizeof(ga));
    ga.mode = GATHER_LANES;
    ga.ph_a = c->ph_a.as<float4>();
    ga.ph_b = c->ph_b.as<float4>();
    ga.off = c->off.as<uint32_t>();
    ga.hp_pos = c->sss_pts.as<float4>();
    ga.hp_w = c->sss_ptw.as<float4>();
    ga.hp_count_src = nullptr;
    ga.n_hit = c->sss_n;
    ga.g = g;
    ga.radius = radius;
    ga.out_rgb = c->sss_E.as<float>();
    launch_gather(s, ga);
    return hipGetLastError();
}

}  // namespace pppmi

extern "C" {

int pppm_subsurface_configure(void* ctx_raw, int64_t points) {
    CHECK_CTX(ctx);
    if (points < 0 || points > (1LL << 26)) return ctx->fail(PPPM_ERR_INVALID, "points: 0 (default) .. 2^26");
    ctx->sss_points_req = points;
    ctx->rendering = false;
    return PPPM_OK;
}

int pppm_subsurface_cut(void* ctx_raw, double max_error) {
    CHECK_CTX(ctx);
    if (!(max_error >= 0.0) || !std::isfinite(max_error))
        return ctx->fail(PPPM_ERR_INVALID, "max_error: 0 (the exact sum) or a positive bound on area / d^2");
    ctx->sss_cut = max_error;
    ctx->rendering = false;
    return PPPM_OK;
}

int pppm_subsurface_points(void* ctx_raw, float* pos, float* normal, int64_t cap, int64_t* n, double* area) {
    CHECK_CTX(ctx);
    if (!ctx->rendering || !ctx->sss_present) return ctx->fail(PPPM_ERR_STATE, "no subsurface points");
    if (n) *n = ctx->sss_n;
    if (area) *area = ctx->sss_area;
    const int64_t m = std::min(cap, ctx->sss_n);
    if (m <= 0) return PPPM_OK;
    (void)hipSetDevice(ctx->device);
    std::vector<float4> p((size_t)ctx->sss_n);
    TRY_HIP(ctx, hipMemcpy(p.data(), ctx->sss_pts.p, sizeof(float4) * p.size(), hipMemcpyDeviceToHost));
    for (int64_t i = 0; i < m; i++) {
        if (pos) {
            pos[3 * i] = p[(size_t)i].x;
            pos[3 * i + 1] = p[(size_t)i].y;
            pos[3 * i + 2] = p[(size_t)i].z;
        }
        if (normal) {
            uint32_t e;
            std::memcpy(&e, &p[(size_t)i].w, 4);
            oct_decode_host(e, normal + 3 * i);
        }
    }
    return PPPM_OK;
}

int pppm_subsurface_irradiance(void* ctx_raw, float* E, int64_t cap, int64_t* n) {
    CHECK_CTX(ctx);
    if (!ctx->last_valid || !ctx->sss_present) return ctx->fail(PPPM_ERR_STATE, "no completed pass with subsurface points");
    if (n) *n = ctx->sss_n;
    const int64_t m = std::min(cap, ctx->sss_n);
    if (m <= 0 || !E) return PPPM_OK;
    (void)hipSetDevice(ctx->device);
    TRY_HIP(ctx, hipMemcpy(E, ctx->sss_E.p, sizeof(float) * 3 * (size_t)m, hipMemcpyDeviceToHost));
    return PPPM_OK;
}

int pppm_last_irradiance_photons(void* ctx_raw, float* pos, float* flux, float* normal, uint32_t* key, int64_t cap,
                                 int64_t* n) {
    CHECK_CTX(ctx);
    if (!ctx->last_valid || !ctx->sss_present) return ctx->fail(PPPM_ERR_STATE, "no completed pass with subsurface points");
    if (n) *n = ctx->last_n_irr;
    const int64_t m = std::min(cap, ctx->last_n_irr);
    if (m <= 0) return PPPM_OK;
    (void)hipSetDevice(ctx->device);
    std::vector<float4> p((size_t)m), f((size_t)m);
    TRY_HIP(ctx, hipMemcpy(p.data(), ctx->irr_pos.p, sizeof(float4) * (size_t)m, hipMemcpyDeviceToHost));
    TRY_HIP(ctx, hipMemcpy(f.data(), ctx->irr_flux.p, sizeof(float4) * (size_t)m, hipMemcpyDeviceToHost));
    for (int64_t i = 0; i < m; i++) {
        if (pos) {
            pos[3 * i] = p[(size_t)i].x;
            pos[3 * i + 1] = p[(size_t)i].y;
            pos[3 * i + 2] = p[(size_t)i].z;
        }
        if (flux) {
            flux[3 * i] = f[(size_t)i].x;
            flux[3 * i + 1] = f[(size_t)i].y;
            flux[3 * i + 2] = f[(size_t)i].z;
        }
        uint32_t e;
        if (normal) {
            std::memcpy(&e, &p[(size_t)i].w, 4);
            oct_decode_host(e, normal + 3 * i);
        }
        if (key) std::memcpy(&key[i], &f[(size_t)i].w, 4);
    }
    return PPPM_OK;
}

int pppm_dipole_eval(void* ctx_raw, const pppm_bsdf* material, const float* pts, const float* E, int64_t n_pts,
                     double area, const float* query, int64_t n_query, float* out) {
    CHECK_CTX(ctx);
    if (!material || !(material->type & PPPM_BSDF_BSSRDF) || n_pts < 0 || n_query < 0 ||
        (n_pts > 0 && (!pts || !E)) || (n_query > 0 && (!query || !out)) || n_pts > (1LL << 26) ||
        n_query > (1LL << 26))
        return ctx->fail(PPPM_ERR_INVALID, "bad dipole arguments");
    if (n_query == 0) return PPPM_OK;
    (void)hipSetDevice(ctx->device);
    const double eta = material->eta > 0.0f ? material->eta : 1.3;
    const double fdr = dipole_fdr(eta), A = (1.0 + fdr) / (1.0 - fdr);
    DipoleDev d;
    std::memset(&d, 0, sizeof(d));
    for (int k = 0; k < 3; k++) {
        const double sa = material->sigma_a[k], ss = material->sigmap_s[k], st = sa + ss;
        const double tr = std::sqrt(3.0 * sa * st);
        d.zp[k] = (float)(1.0 / st);
        d.zn[k] = (float)((1.0 / st) * (1.0 + (4.0 / 3.0) * A));
        d.tr[k] = (float)tr;
        d.k[k] = (float)((ss / st) / (4.0 * M_PI * tr));
    }
    d.out = (float)((1.0 - fdr) / M_PI);
    std::vector<float4> p4((size_t)std::max<int64_t>(n_pts, 1)), q4((size_t)n_query), w4((size_t)n_query);
    for (int64_t i = 0; i < n_pts; i++) p4[(size_t)i] = make_float4(pts[3 * i], pts[3 * i + 1], pts[3 * i + 2], 0.0f);
    for (int64_t i = 0; i < n_query; i++) {
        q4[(size_t)i] = make_float4(query[3 * i], query[3 * i + 1], query[3 * i + 2], 0.0f);  // dipole index 0
        const uint32_t id = (uint32_t)i;
        float idf;
        std::memcpy(&idf, &id, 4);
        w4[(size_t)i] = make_float4(1.0f, 1.0f, 1.0f, idf);
    }
    CutBufs cut;  // freed on return
    void* buf = nullptr;
    const size_t bytes = sizeof(DipoleDev) + sizeof(float4) * (p4.size() + 2 * q4.size()) +
                         sizeof(float) * 3 * (size_t)(std::max<int64_t>(n_pts, 1) + n_query);
    TRY_HIP(ctx, hipMalloc(&buf, bytes));
    char* c = static_cast<char*>(buf);
    DipoleDev* dd = reinterpret_cast<DipoleDev*>(c);
    float4* dp = reinterpret_cast<float4*>(c + sizeof(DipoleDev));
    float4* dq = dp + p4.size();
    float4* dw = dq + q4.size();
    float* dE = reinterpret_cast<float*>(dw + q4.size());
    float* dout = dE + 3 * std::max<int64_t>(n_pts, 1);
    hipError_t e = hipMemcpy(dd, &d, sizeof(d), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(dp, p4.data(), sizeof(float4) * p4.size(), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(dq, q4.data(), sizeof(float4) * q4.size(), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(dw, w4.data(), sizeof(float4) * w4.size(), hipMemcpyHostToDevice);
    if (e == hipSuccess && n_pts > 0) e = hipMemcpy(dE, E, sizeof(float) * 3 * (size_t)n_pts, hipMemcpyHostToDevice);
    if (e == hipSuccess) {
        DipoleArgs a;
        a.pts = dp;
        a.E = dE;
        a.n_pts = n_pts;
        a.area = (float)area;
        a.dip = dd;
        a.hit_pos = dq;
        a.hit_w = dw;
        a.count_src = nullptr;
        a.n_hit = n_query;
        a.img_sum = nullptr;
        a.img_comp = nullptr;
        a.out_rgb = dout;
        if (n_pts > 0 && ctx->sss_cut > 0.0) {  // the cut over these points, built for this call
            CutHost h;
            cut_build(p4.data(), n_pts, area, &h);
            e = cut_upload(&cut, h);
            if (e == hipSuccess) {
                cut_refresh(cut, ctx->stream, dE, ctx->sss_cut);
                launch_dipole_cut(ctx->stream, a, cut_dev(cut, ctx->sss_cut));
            }
        } else if (n_pts > 0) {
            launch_dipole(ctx->stream, a);
        } else {
            launch_zero_u32(ctx->stream, dout, 3 * n_query);
        }
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpyAsync(out, dout, sizeof(float) * 3 * (size_t)n_query, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    (void)hipFree(buf);
    if (e != hipSuccess) return ctx->fail(PPPM_ERR_HIP, std::string("dipole: ") + hipGetErrorString(e));
    return PPPM_OK;
}

}  // extern "C"
