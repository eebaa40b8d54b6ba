// capi.hip -- the C ABI of libpppm_hip.so (declared in include/pppm.h):
// context, scene upload, BVH, batch intersection, hash grid, radius
// schedule, camera, image read-back and the inspection / measurement hooks.
// The pass loop of ProgressivePhotonMappingProb::render
// (sources/ppm_probabilistic.cc:24-121) lives in pass.hip, subsurface
// scattering in sss_host.hip, the learned proposal in proposal_host.hip and
// batched scenes in batch.hip.
#include "ctx.h"
#include "../host/image_io.h"

using namespace pppmi;

namespace pppmi {

int bits_for(uint64_t max_value) {
    int b = 0;
    while (b < 64 && (max_value >> b) != 0) b++;
    return b;
}

// host copy of the device octahedral encode (pppm_device.h), returned as the
// float whose bits hold the code
float oct_encode_host(v3f n) {
    const float s = std::fabs(n.x) + std::fabs(n.y) + std::fabs(n.z);
    float x = n.x / s, y = n.y / s;
    if (n.z < 0.0f) {
        const float ox = (1.0f - std::fabs(y)) * (x < 0.0f ? -1.0f : 1.0f);
        const float oy = (1.0f - std::fabs(x)) * (y < 0.0f ? -1.0f : 1.0f);
        x = ox;
        y = oy;
    }
    const int qx = (int)std::rint(std::fmin(std::fmax(x, -1.0f), 1.0f) * 32767.0f);
    const int qy = (int)std::rint(std::fmin(std::fmax(y, -1.0f), 1.0f) * 32767.0f);
    const uint32_t e = ((uint32_t)(qx & 0xFFFF)) | ((uint32_t)(qy & 0xFFFF) << 16);
    float f;
    std::memcpy(&f, &e, 4);
    return f;
}

// host copy of the device octahedral decode (pppm_device.h)
void oct_decode_host(uint32_t e, float out[3]) {
    float x = (float)(int16_t)(e & 0xFFFF) * (1.0f / 32767.0f);
    float y = (float)(int16_t)(e >> 16) * (1.0f / 32767.0f);
    float z = 1.0f - std::fabs(x) - std::fabs(y);
    if (z < 0.0f) {
        float ox = (1.0f - std::fabs(y)) * (x < 0.0f ? -1.0f : 1.0f);
        float oy = (1.0f - std::fabs(x)) * (y < 0.0f ? -1.0f : 1.0f);
        x = ox;
        y = oy;
    }
    float n = std::sqrt(x * x + y * y + z * z);
    out[0] = x / n;
    out[1] = y / n;
    out[2] = z / n;
}

void cross3(const double a[3], const double b[3], double o[3]) {
    o[0] = a[1] * b[2] - a[2] * b[1];
    o[1] = a[2] * b[0] - a[0] * b[2];
    o[2] = a[0] * b[1] - a[1] * b[0];
}

// Builds the grid (sort by bucket of (pos4 or pos3) through perm) into the
// context's sort buffers, avoiding the buffer `keep` (index into u[], or -1).
// Returns the buffer holding the sorted values (grid sorted_idx).
// With a device count the n slots hold *count records followed by unused
// ones, whose keys sort behind every bucket.
// Bits of a bucket id in the grid sort key (with a device count, one more
// value for the sentinel of unused slots).
int grid_bucket_bits(uint32_t size, bool sentinel) {
    return sentinel ? bits_for((uint64_t)size) : bits_for(size > 0 ? (uint64_t)(size - 1) : 0);
}

// Cell-local Morton bits per axis below the bucket in the render path's grid
// sort (at most 3, keeping the key within 32 bits): the lanes gather walks a
// bucket in blocks whose boxes are then small.  Other gather kernels walk whole
// buckets and keep the reference's insertion order.
// The bucket bits always count a sentinel value, so single passes and
// enqueued (batched) passes order a bucket's records the same way.
int render_sub_bits(const Ctx* c, uint32_t size) {
    if (!gather_compact(c->rp.gather_mode)) return 0;
    return std::max(0, std::min(3, (32 - grid_bucket_bits(size, true)) / 3));
}

uint32_t* build_grid(Ctx* c, const float4* pos4, const float* pos3, const uint32_t* perm, int64_t n,
                     const GridParams& g, int keep, const unsigned long long* count, uint32_t* off_out,
                     bool keys_ready) {
    uint32_t* b[4];
    int k = 0;
    for (int i = 0; i < 5 && k < 4; i++)
        if (i != keep) b[k++] = c->u[i].as<uint32_t>();
    const int bucket_bits = grid_bucket_bits(g.size, count != nullptr);
    const int shift = 3 * g.sub_bits;
    const int bits = bucket_bits + shift;
    const uint32_t sentinel = bits >= 32 ? 0xFFFFFFFFu : (uint32_t)((1ull << bits) - 1);
    if (!keys_ready) launch_point_keys(c->stream, pos4, pos3, perm, n, g, b[0], b[1], count, sentinel);
    bool alt = radix_sort_pairs(c->stream, b[0], b[1], b[2], b[3], c->hist.as<uint32_t>(), n, bits);
    uint32_t* keys = alt ? b[2] : b[0];
    uint32_t* vals = alt ? b[3] : b[1];
    // the radix histogram is free again: scratch of the offsets' suffix minimum
    launch_bucket_offsets(c->stream, keys, n, g.size, off_out ? off_out : c->off.as<uint32_t>(), shift,
                          c->hist.as<uint32_t>(), (int64_t)(c->hist.n / sizeof(uint32_t)));
    return vals;
}

// The photon records of a pass in origin order (stable by photon index) with
// their grid keys, from the ranks and counts the photon pass wrote: perm1
// (origin position -> record slot) in u[4], keys and values (record slots) in
// u[0], u[1] for build_grid(..., keep = 4, keys_ready).  Returns perm1.
void origin_starts(Ctx* c, hipStream_t s) {
    launch_origin_starts(s, c->rec_count.as<uint32_t>(), c->rp.photons, c->rec_tiles.as<uint32_t>(),
                         c->rec_start.as<uint32_t>());
}

uint32_t* origin_order(Ctx* c, hipStream_t s, int64_t n, const GridParams& g, const unsigned long long* count,
                       bool starts_done) {
    const int bits = grid_bucket_bits(g.size, count != nullptr) + 3 * g.sub_bits;
    const uint32_t sentinel = bits >= 32 ? 0xFFFFFFFFu : (uint32_t)((1ull << bits) - 1);
    if (!starts_done) origin_starts(c, s);
    launch_origin_point_keys(s, c->ph_rec.as<float4>(), c->ph_rec.as<float4>() + 1, c->rec_rank.as<uint8_t>(),
                             c->rec_start.as<uint32_t>(), n, count, g, sentinel, c->u[0].as<uint32_t>(),
                             c->u[1].as<uint32_t>(), c->u[4].as<uint32_t>(), 2);
    return c->u[4].as<uint32_t>();
}

hipError_t ensure_sort_buffers(Ctx* c, int64_t n, uint32_t S) {
    for (int i = 0; i < 5; i++) {
        hipError_t e = c->u[i].ensure(sizeof(uint32_t) * (size_t)(n > 0 ? n : 1));
        if (e != hipSuccess) return e;
    }
    hipError_t e = c->hist.ensure(sizeof(uint32_t) * (size_t)radix_hist_words(n > 0 ? n : 1));
    if (e != hipSuccess) return e;
    return c->off.ensure(sizeof(uint32_t) * ((size_t)S + 1));
}

hipError_t ensure_hit_sort(Ctx* c, int64_t n) {
    const size_t m = (size_t)(n > 0 ? n : 1);
    hipError_t e;
    for (auto* b : {&c->hp_pos_s, &c->hp_w_s})
        if ((e = b->ensure(sizeof(float4) * m)) != hipSuccess) return e;
    for (auto& b : c->hk)
        if ((e = b.ensure(sizeof(uint32_t) * m)) != hipSuccess) return e;
    return c->hhist.ensure(sizeof(uint32_t) * (size_t)radix_hist_words((int64_t)m));
}

}  // namespace pppmi

namespace {
thread_local std::string g_create_error;
}  // namespace

extern "C" {

int pppm_create(void** out, int32_t device) {
    if (!out) return PPPM_ERR_INVALID;
    *out = nullptr;
    int count = 0;
    hipError_t e = hipGetDeviceCount(&count);
    if (e != hipSuccess || count == 0) {
        g_create_error = std::string("no HIP device: ") + hipGetErrorString(e);
        return PPPM_ERR_HIP;
    }
    if (device < 0 || device >= count) {
        g_create_error = "device index out of range";
        return PPPM_ERR_INVALID;
    }
    Ctx* c = new Ctx();
    c->device = device;
    (void)hipSetDevice(device);
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) == hipSuccess) {
        c->name = prop.gcnArchName;
        c->cu = prop.multiProcessorCount;
        c->hbm = prop.totalGlobalMem;
    }
    c->noalloc_err = noalloc_setup(device, &c->noalloc_scratch);
    if (hipStreamCreate(&c->stream) != hipSuccess) {
        g_create_error = "hipStreamCreate failed";
        delete c;
        return PPPM_ERR_HIP;
    }
    if (hipStreamCreate(&c->stream2) != hipSuccess || hipStreamCreate(&c->stream3) != hipSuccess ||
        hipStreamCreate(&c->stream4) != hipSuccess) {
        g_create_error = "hipStreamCreate failed";
        if (c->stream2) (void)hipStreamDestroy(c->stream2);
        if (c->stream3) (void)hipStreamDestroy(c->stream3);
        (void)hipStreamDestroy(c->stream);
        delete c;
        return PPPM_ERR_HIP;
    }
    for (auto& ev : c->ev) (void)hipEventCreate(&ev);
    for (auto& set : c->pev)
        for (auto& ev : set) (void)hipEventCreate(&ev);
    *out = c;
    return PPPM_OK;
}

int pppm_destroy(void* ctx_raw) {
    Ctx* c = reinterpret_cast<Ctx*>(ctx_raw);
    if (!c) return PPPM_ERR_INVALID;
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
    c->tri9.release(); c->tri_mat.release(); c->mats.release(); c->lights.release();
    (void)hipFree(c->nodes); (void)hipFree(c->tris4);
    c->ph_rec.release(); c->ph_a.release(); c->ph_b.release(); c->ph_c.release();
    for (auto& b : c->u) b.release();
    c->hist.release(); c->off.release(); c->hp_pos.release(); c->hp_w.release();
    c->img_sum.release(); c->img_comp.release(); c->img_dsum.release(); c->img_dcomp.release();
    c->counters.release(); c->scratch.release();
    c->rec_rank.release(); c->rec_count.release(); c->rec_start.release(); c->rec_tiles.release();
    (void)hipStreamSynchronize(c->stream2);
    (void)hipStreamSynchronize(c->stream3);
    (void)hipStreamSynchronize(c->stream4);
    for (auto& ev : c->ev) (void)hipEventDestroy(ev);
    for (auto& set : c->pev)
        for (auto& ev : set) (void)hipEventDestroy(ev);
    (void)hipStreamDestroy(c->stream4);
    (void)hipStreamDestroy(c->stream3);
    (void)hipStreamDestroy(c->stream2);
    (void)hipStreamDestroy(c->stream);
    (void)hipHostFree(c->pinned_counters);
    delete c;
    return PPPM_OK;
}

const char* pppm_last_error(const void* ctx_raw) {
    const Ctx* c = reinterpret_cast<const Ctx*>(ctx_raw);
    if (!c) return g_create_error.c_str();
    return c->err.c_str();
}

int pppm_device_info(void* ctx_raw, char* name, int32_t name_len, int32_t* cu_count,
                     int64_t* hbm_bytes) {
    CHECK_CTX(ctx);
    if (name && name_len > 0) {
        std::snprintf(name, (size_t)name_len, "%s", ctx->name.c_str());
    }
    if (cu_count) *cu_count = ctx->cu;
    if (hbm_bytes) *hbm_bytes = (int64_t)ctx->hbm;
    return PPPM_OK;
}

int pppm_scene_upload(void* ctx_raw, const float* tri_xyz, const int32_t* bsdf_id, int64_t n,
                      const pppm_bsdf* table, int32_t nb, const pppm_light* lights, int32_t nl) {
    CHECK_CTX(ctx);
    if (n < 0 || (n > 0 && (!tri_xyz || !bsdf_id)) || nb <= 0 || !table || nl < 0 || (nl > 0 && !lights))
        return ctx->fail(PPPM_ERR_INVALID, "scene needs triangles, bsdf ids and a bsdf table");
    for (int32_t i = 0; i < nb; i++) {
        int t = table[i].type;
        if (t & PPPM_BSDF_PHONG) return ctx->fail(PPPM_ERR_INVALID, "Phong BRDF is not supported");
        if ((t & PPPM_BSDF_BSSRDF) &&
            (!(t & PPPM_BSDF_REFRACTION) || !(table[i].sigma_a[0] > 0.0f && table[i].sigma_a[1] > 0.0f &&
                                              table[i].sigma_a[2] > 0.0f && table[i].sigmap_s[0] > 0.0f &&
                                              table[i].sigmap_s[1] > 0.0f && table[i].sigmap_s[2] > 0.0f)))
            return ctx->fail(PPPM_ERR_INVALID, "a BSSRDF needs the refraction boundary and positive sigma_a, sigma'_s");
        if (!(t & (PPPM_BSDF_LAMBERTIAN | PPPM_BSDF_SPECULAR | PPPM_BSDF_REFRACTION | PPPM_BSDF_EMITTER)))
            return ctx->fail(PPPM_ERR_INVALID, "bsdf " + std::to_string(i) + " has no supported type");
    }
    for (int64_t i = 0; i < n; i++)
        if (bsdf_id[i] < 0 || bsdf_id[i] >= nb)
            return ctx->fail(PPPM_ERR_INVALID, "bsdf id out of range at triangle " + std::to_string(i));
    // the BVH builders order boxes by Morton codes and surface areas, which a
    // NaN or infinite coordinate leaves undefined
    for (int64_t i = 0; i < 9 * n; i++)
        if (!std::isfinite(tri_xyz[i]))
            return ctx->fail(PPPM_ERR_INVALID, "non-finite coordinate at triangle " + std::to_string(i / 9));
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    ctx->has_scene = false;
    ctx->has_bvh = false;
    ctx->rendering = false;
    TRY_HIP(ctx, ctx->tri9.ensure(sizeof(float) * 9 * (size_t)n));
    TRY_HIP(ctx, ctx->tri_mat.ensure(sizeof(int32_t) * (size_t)n));
    if (n > 0) {
        TRY_HIP(ctx, hipMemcpy(ctx->tri9.p, tri_xyz, sizeof(float) * 9 * (size_t)n, hipMemcpyHostToDevice));
        TRY_HIP(ctx, hipMemcpy(ctx->tri_mat.p, bsdf_id, sizeof(int32_t) * (size_t)n, hipMemcpyHostToDevice));
    }
    ctx->h_bsdf.assign(table, table + nb);
    ctx->h_lights.assign(lights, lights + nl);
    ctx->has_env = false;  // an environment map belongs to the scene: set it after the upload
    std::vector<Mat> m(nb);
    for (int32_t i = 0; i < nb; i++) {
        const pppm_bsdf& b = table[i];
        std::memset(&m[i], 0, sizeof(Mat));
        m[i].type = b.type;
        for (int k = 0; k < 3; k++) {
            m[i].rho[k] = b.reflectance[k];
            m[i].le[k] = b.emission[k];
        }
        m[i].rr_mean = (float)(((double)b.reflectance[0] + (double)b.reflectance[1] + (double)b.reflectance[2]) / 3.0);
        m[i].rr_max = std::fmax(b.reflectance[0], std::fmax(b.reflectance[1], b.reflectance[2]));
        m[i].ior = b.ior > 0.0f ? b.ior : 1.5f;
        m[i].sss = -1;
    }
    // dipole constants of the BSSRDF materials (sources/bssrdf.cc:33-62, FP64)
    std::vector<DipoleDev> dv;
    ctx->h_sss_index.assign((size_t)nb, -1);
    for (int32_t i = 0; i < nb; i++) {
        const pppm_bsdf& b = table[i];
        if (!(b.type & PPPM_BSDF_BSSRDF)) continue;
        const double eta = b.eta > 0.0f ? b.eta : 1.3;
        DipoleDev d;
        std::memset(&d, 0, sizeof(d));
        const double fdr = dipole_fdr(eta);
        const double A = (1.0 + fdr) / (1.0 - fdr);
        for (int k = 0; k < 3; k++) {
            const double sa = b.sigma_a[k], ss = b.sigmap_s[k];
            const double st = sa + ss;
            const double tr = std::sqrt(3.0 * sa * st);
            d.zp[k] = (float)(1.0 / st);
            d.zn[k] = (float)((1.0 / st) * (1.0 + (4.0 / 3.0) * A));
            d.tr[k] = (float)tr;
            d.k[k] = (float)((ss / st) / (4.0 * M_PI * tr));
        }
        d.out = (float)((1.0 - fdr) / M_PI);
        m[i].sss = (int32_t)dv.size();
        ctx->h_sss_index[(size_t)i] = m[i].sss;
        dv.push_back(d);
    }
    ctx->n_dips = (int32_t)dv.size();
    if (!dv.empty()) {
        TRY_HIP(ctx, ctx->dips.ensure(sizeof(DipoleDev) * dv.size()));
        TRY_HIP(ctx, hipMemcpy(ctx->dips.p, dv.data(), sizeof(DipoleDev) * dv.size(), hipMemcpyHostToDevice));
    }
    TRY_HIP(ctx, ctx->mats.ensure(sizeof(Mat) * (size_t)nb));
    TRY_HIP(ctx, hipMemcpy(ctx->mats.p, m.data(), sizeof(Mat) * (size_t)nb, hipMemcpyHostToDevice));
    TRY_HIP(ctx, ctx->lights.ensure(sizeof(LightDev) * (size_t)nl));
    for (int d = 0; d < 3; d++) {
        ctx->bmin[d] = 1e300;
        ctx->bmax[d] = -1e300;
    }
    for (int64_t i = 0; i < n; i++)
        for (int j = 0; j < 9; j++) {
            double v = tri_xyz[9 * i + j];
            ctx->bmin[j % 3] = std::fmin(ctx->bmin[j % 3], v);
            ctx->bmax[j % 3] = std::fmax(ctx->bmax[j % 3], v);
        }
    ctx->n_tris = n;
    ctx->has_scene = true;
    return PPPM_OK;
}

int pppm_bvh_build(void* ctx_raw, pppm_bvh_info* info) { return pppm_bvh_build_with(ctx_raw, PPPM_BVH_DEFAULT, info); }

int pppm_bvh_build_with(void* ctx_raw, int32_t builder, pppm_bvh_info* info) {
    CHECK_CTX(ctx);
    if (!ctx->has_scene) return ctx->fail(PPPM_ERR_STATE, "no scene uploaded");
    if (builder != PPPM_BVH_DEFAULT && builder != PPPM_BVH_LBVH && builder != PPPM_BVH_PLOC)
        return ctx->fail(PPPM_ERR_INVALID, "unknown BVH builder");
    if (builder == PPPM_BVH_DEFAULT) builder = BVH_BUILDER_DEFAULT;
    (void)hipSetDevice(ctx->device);
    (void)hipFree(ctx->nodes);
    (void)hipFree(ctx->tris4);
    ctx->nodes = nullptr;
    ctx->tris4 = nullptr;
    ctx->has_bvh = false;
    TRY_HIP(ctx, hipEventRecord(ctx->ev[0], ctx->stream));
    BvhBuild b;
    std::string e = bvh_build(ctx->stream, ctx->tri9.as<float>(), ctx->n_tris,
                              builder == PPPM_BVH_PLOC ? BVH_BUILD_PLOC : BVH_BUILD_LBVH, &b);
    if (!e.empty()) return ctx->fail(PPPM_ERR_HIP, "BVH build: " + e);
    if (builder == PPPM_BVH_PLOC && b.builder_used != BVH_BUILD_PLOC) builder = PPPM_BVH_LBVH;
    // the traversal stack holds at most 3 entries per level below the root:
    // a clustered tree deeper than that is replaced by the radix tree, whose
    // depth is bounded by the 30 Morton bits
    if (builder == PPPM_BVH_PLOC && 3 * (int64_t)b.depth + 1 > STACK_MAX) {
        (void)hipFree(b.nodes);
        (void)hipFree(b.tris4);
        builder = PPPM_BVH_LBVH;
        e = bvh_build(ctx->stream, ctx->tri9.as<float>(), ctx->n_tris, BVH_BUILD_LBVH, &b);
        if (!e.empty()) return ctx->fail(PPPM_ERR_HIP, "BVH build: " + e);
    }
    TRY_HIP(ctx, hipEventRecord(ctx->ev[1], ctx->stream));
    TRY_HIP(ctx, hipEventSynchronize(ctx->ev[1]));
    float ms = 0.0f;
    (void)hipEventElapsedTime(&ms, ctx->ev[0], ctx->ev[1]);
    ctx->nodes = b.nodes;
    ctx->tris4 = b.tris4;
    ctx->root = b.root;
    if (3 * (int64_t)b.depth + 1 > STACK_MAX)
        return ctx->fail(PPPM_ERR_INVALID, "BVH deeper than the traversal stack allows");
    // node fetches address the node array with 32-bit byte offsets
    if ((int64_t)b.n_nodes * (int64_t)sizeof(Node4) > (int64_t)INT32_MAX)
        return ctx->fail(PPPM_ERR_INVALID, "BVH larger than 2 GiB of nodes");
    ctx->binfo.nodes = b.n_nodes;
    ctx->binfo.leaves = b.n_leaves;
    ctx->binfo.depth = b.depth;
    ctx->binfo.max_leaf = LEAF_MAX;
    ctx->binfo.build_ms = ms;
    ctx->binfo.builder = builder;
    ctx->binfo.reserved = 0;
    ctx->has_bvh = true;
    if (info) *info = ctx->binfo;
    return PPPM_OK;
}

int pppm_intersect(void* ctx_raw, const pppm_ray* rays, int64_t n, pppm_hit* hits) {
    CHECK_CTX(ctx);
    if (!ctx->has_bvh) return ctx->fail(PPPM_ERR_STATE, "no BVH built");
    if (n < 0 || (n > 0 && (!rays || !hits))) return ctx->fail(PPPM_ERR_INVALID, "bad ray arrays");
    if (n == 0) return PPPM_OK;
    (void)hipSetDevice(ctx->device);
    // placeholder lights: intersection needs none, but SceneDev reads n_lights only
    TRY_HIP(ctx, ctx->scratch.ensure((sizeof(pppm_ray) + sizeof(pppm_hit)) * (size_t)n));
    pppm_ray* dr = ctx->scratch.as<pppm_ray>();
    pppm_hit* dh = reinterpret_cast<pppm_hit*>(dr + n);
    TRY_HIP(ctx, hipMemcpyAsync(dr, rays, sizeof(pppm_ray) * (size_t)n, hipMemcpyHostToDevice, ctx->stream));
    launch_intersect(ctx->stream, ctx->scene_dev(), dr, n, dh);
    TRY_HIP(ctx, hipGetLastError());
    TRY_HIP(ctx, hipMemcpyAsync(hits, dh, sizeof(pppm_hit) * (size_t)n, hipMemcpyDeviceToHost, ctx->stream));
    TRY_HIP(ctx, hipStreamSynchronize(ctx->stream));
    for (int64_t i = 0; i < n; i++)
        if (hits[i].tri == -2) return ctx->fail(PPPM_ERR_HIP, "traversal stack overflow");
    return PPPM_OK;
}

int pppm_bvh_download(void* ctx_raw, pppm_bvh_node* nodes, int64_t node_cap, float* tris, int64_t tri_cap,
                      int64_t* n_nodes, int64_t* n_tris, int32_t* root) {
    CHECK_CTX(ctx);
    if (!ctx->has_bvh) return ctx->fail(PPPM_ERR_STATE, "no BVH built");
    static_assert(sizeof(pppm_bvh_node) == sizeof(Node4), "pppm_bvh_node mirrors the device node");
    const int64_t nn = ctx->binfo.nodes, nt = ctx->n_tris;
    if (n_nodes) *n_nodes = nn;
    if (n_tris) *n_tris = nt;
    if (root) *root = ctx->root;
    if ((nodes && node_cap < nn) || (tris && tri_cap < nt))
        return ctx->fail(PPPM_ERR_CAPACITY, "bvh download: buffers smaller than the tree");
    (void)hipSetDevice(ctx->device);
    if (nodes && nn > 0)
        TRY_HIP(ctx, hipMemcpyAsync(nodes, ctx->nodes, sizeof(Node4) * (size_t)nn, hipMemcpyDeviceToHost, ctx->stream));
    if (tris && nt > 0)
        TRY_HIP(ctx, hipMemcpyAsync(tris, ctx->tris4, sizeof(float4) * 3 * (size_t)nt, hipMemcpyDeviceToHost,
                                    ctx->stream));
    TRY_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return PPPM_OK;
}

int pppm_hashgrid_build(void* ctx_raw, const float* pos, int64_t n, int32_t hash_size,
                        double hash_scale, const double bbox_min[3], int32_t* sorted_idx,
                        int32_t* cell_start, int32_t* cell_end) {
    CHECK_CTX(ctx);
    if (n < 0 || (n > 0 && !pos) || hash_size <= 0 || !bbox_min || n > 0x7FFFFFFFLL)
        return ctx->fail(PPPM_ERR_INVALID, "bad hash grid arguments");
    (void)hipSetDevice(ctx->device);
    TRY_HIP(ctx, ensure_sort_buffers(ctx, n, (uint32_t)hash_size));
    TRY_HIP(ctx, ctx->scratch.ensure(sizeof(float) * 3 * (size_t)(n > 0 ? n : 1)));
    if (n > 0)
        TRY_HIP(ctx, hipMemcpyAsync(ctx->scratch.p, pos, sizeof(float) * 3 * (size_t)n,
                                    hipMemcpyHostToDevice, ctx->stream));
    GridParams g;
    for (int d = 0; d < 3; d++) g.bmin[d] = bbox_min[d];
    g.scale = hash_scale;
    g.size = (uint32_t)hash_size;
    uint32_t* vals = build_grid(ctx, nullptr, ctx->scratch.as<float>(), nullptr, n, g, -1);
    TRY_HIP(ctx, hipGetLastError());
    std::vector<uint32_t> off((size_t)hash_size + 1);
    if (n > 0 && sorted_idx)
        TRY_HIP(ctx, hipMemcpyAsync(sorted_idx, vals, sizeof(uint32_t) * (size_t)n, hipMemcpyDeviceToHost, ctx->stream));
    TRY_HIP(ctx, hipMemcpyAsync(off.data(), ctx->off.p, sizeof(uint32_t) * off.size(), hipMemcpyDeviceToHost, ctx->stream));
    TRY_HIP(ctx, hipStreamSynchronize(ctx->stream));
    copy_ranges(off, hash_size, cell_start, cell_end);
    return PPPM_OK;
}

int pppm_hashgrid_build_boxes(void* ctx_raw, const double* box_min, const double* box_max, int64_t n,
                              int32_t hash_size, double hash_scale, const double bbox_min[3],
                              int32_t* sorted_items, int64_t cap, int64_t* n_pairs,
                              int32_t* cell_start, int32_t* cell_end) {
    CHECK_CTX(ctx);
    if (n < 0 || (n > 0 && (!box_min || !box_max)) || hash_size <= 0 || !bbox_min || !n_pairs)
        return ctx->fail(PPPM_ERR_INVALID, "bad hash grid arguments");
    (void)hipSetDevice(ctx->device);
    GridParams g;
    for (int d = 0; d < 3; d++) g.bmin[d] = bbox_min[d];
    g.scale = hash_scale;
    g.size = (uint32_t)hash_size;
    int64_t nn = n > 0 ? n : 1;
    int64_t scan_words = nn + nn / 1024 + 64;
    TRY_HIP(ctx, ctx->scratch.ensure(sizeof(double) * 6 * (size_t)nn + sizeof(uint64_t) * (size_t)(nn + scan_words)));
    double* dmin = ctx->scratch.as<double>();
    double* dmax = dmin + 3 * nn;
    uint64_t* counts = reinterpret_cast<uint64_t*>(dmax + 3 * nn);
    uint64_t* bsum = counts + nn;
    if (n > 0) {
        TRY_HIP(ctx, hipMemcpyAsync(dmin, box_min, sizeof(double) * 3 * (size_t)n, hipMemcpyHostToDevice, ctx->stream));
        TRY_HIP(ctx, hipMemcpyAsync(dmax, box_max, sizeof(double) * 3 * (size_t)n, hipMemcpyHostToDevice, ctx->stream));
    }
    launch_box_counts(ctx->stream, dmin, dmax, n, g, counts);
    uint64_t last_count = 0;
    if (n > 0)
        TRY_HIP(ctx, hipMemcpyAsync(&last_count, counts + (n - 1), sizeof(uint64_t), hipMemcpyDeviceToHost, ctx->stream));
    TRY_HIP(ctx, hipStreamSynchronize(ctx->stream));
    launch_exclusive_scan_u64(ctx->stream, counts, n, bsum);
    uint64_t last_off = 0;
    if (n > 0)
        TRY_HIP(ctx, hipMemcpyAsync(&last_off, counts + (n - 1), sizeof(uint64_t), hipMemcpyDeviceToHost, ctx->stream));
    TRY_HIP(ctx, hipStreamSynchronize(ctx->stream));
    int64_t total = (int64_t)(last_off + last_count);
    *n_pairs = total;
    if (total > cap) return ctx->fail(PPPM_ERR_CAPACITY, "pair count exceeds cap");
    if (total > 0x7FFFFFFFLL) return ctx->fail(PPPM_ERR_INVALID, "too many (item, cell) pairs");
    TRY_HIP(ctx, ensure_sort_buffers(ctx, total, (uint32_t)hash_size));
    uint32_t* b[4] = {ctx->u[0].as<uint32_t>(), ctx->u[1].as<uint32_t>(), ctx->u[2].as<uint32_t>(),
                      ctx->u[3].as<uint32_t>()};
    launch_box_pairs(ctx->stream, dmin, dmax, n, g, counts, b[0], b[1]);
    bool alt = radix_sort_pairs(ctx->stream, b[0], b[1], b[2], b[3], ctx->hist.as<uint32_t>(), total,
                                bits_for((uint64_t)(hash_size - 1)));
    uint32_t* keys = alt ? b[2] : b[0];
    uint32_t* vals = alt ? b[3] : b[1];
    launch_bucket_offsets(ctx->stream, keys, total, g.size, ctx->off.as<uint32_t>());
    TRY_HIP(ctx, hipGetLastError());
    std::vector<uint32_t> off((size_t)hash_size + 1);
    if (total > 0 && sorted_items)
        TRY_HIP(ctx, hipMemcpyAsync(sorted_items, vals, sizeof(uint32_t) * (size_t)total, hipMemcpyDeviceToHost, ctx->stream));
    TRY_HIP(ctx, hipMemcpyAsync(off.data(), ctx->off.p, sizeof(uint32_t) * off.size(), hipMemcpyDeviceToHost, ctx->stream));
    TRY_HIP(ctx, hipStreamSynchronize(ctx->stream));
    copy_ranges(off, hash_size, cell_start, cell_end);
    return PPPM_OK;
}

int pppm_radius_schedule(double r1, double alpha, int32_t n, int32_t literal, double* out) {
    if (n < 0 || (n > 0 && !out) || !(r1 > 0.0) || !(alpha > 0.0)) return PPPM_ERR_INVALID;
    if (n == 0) return PPPM_OK;
    out[0] = r1;
    if (literal) {
        double r = r1;
        for (int t = 1; t < n; t++) {
            r = (t + 1.0) / (t + alpha) * r;  // ppm_probabilistic.cc:94
            out[t] = r;
        }
    } else {
        double r2 = r1 * r1;
        for (int i = 1; i < n; i++) {
            r2 = r2 * ((i + alpha) / (i + 1.0));  // Knaus & Zwicker 2011
            out[i] = std::sqrt(r2);
        }
    }
    return PPPM_OK;
}

int pppm_camera_set(void* ctx_raw, const pppm_camera* cam) {
    CHECK_CTX(ctx);
    if (!cam || cam->width <= 0 || cam->height <= 0 || !(cam->fov_y_deg > 0.0 && cam->fov_y_deg < 180.0))
        return ctx->fail(PPPM_ERR_INVALID, "bad camera");
    // pixel ids travel as 32-bit words, all ones meaning none
    if ((int64_t)cam->width * (int64_t)cam->height > (int64_t)1 << 31)
        return ctx->fail(PPPM_ERR_INVALID, "at most 2^31 pixels");
    ctx->cam = *cam;
    ctx->has_cam = true;
    ctx->rendering = false;
    return PPPM_OK;
}

int pppm_image_read(void* ctx_raw, double* rgb, int64_t n_values) {
    CHECK_CTX(ctx);
    if (!ctx->rendering) return ctx->fail(PPPM_ERR_STATE, "no render in progress");
    const int64_t nv = 3 * (int64_t)ctx->cam.width * ctx->cam.height;
    if (!rgb || n_values < nv) return ctx->fail(PPPM_ERR_INVALID, "image buffer too small");
    (void)hipSetDevice(ctx->device);
    if (ctx->rp.method == PPPM_METHOD_HACHISUKA) {
        TRY_HIP(ctx, ctx->ppm_image.ensure(sizeof(double) * (size_t)nv));
        TRY_HIP(ctx, hachisuka_image(ctx, ctx->ppm_image.as<double>()));
        TRY_HIP(ctx, hipMemcpyAsync(rgb, ctx->ppm_image.p, sizeof(double) * (size_t)nv, hipMemcpyDeviceToHost, ctx->stream));
        TRY_HIP(ctx, hipStreamSynchronize(ctx->stream));
        return PPPM_OK;
    }
    std::vector<float> s((size_t)nv), c((size_t)nv), ds((size_t)nv), dc((size_t)nv);
    TRY_HIP(ctx, hipStreamSynchronize(ctx->stream2));
    TRY_HIP(ctx, hipMemcpyAsync(s.data(), ctx->img_sum.p, sizeof(float) * (size_t)nv, hipMemcpyDeviceToHost, ctx->stream));
    TRY_HIP(ctx, hipMemcpyAsync(c.data(), ctx->img_comp.p, sizeof(float) * (size_t)nv, hipMemcpyDeviceToHost, ctx->stream));
    TRY_HIP(ctx, hipMemcpyAsync(ds.data(), ctx->img_dsum.p, sizeof(float) * (size_t)nv, hipMemcpyDeviceToHost, ctx->stream));
    TRY_HIP(ctx, hipMemcpyAsync(dc.data(), ctx->img_dcomp.p, sizeof(float) * (size_t)nv, hipMemcpyDeviceToHost, ctx->stream));
    TRY_HIP(ctx, hipStreamSynchronize(ctx->stream));
    // FP64: the compensated sums of both accumulators, averaged over the passes
    const double inv = ctx->passes_done > 0 ? 1.0 / (double)ctx->passes_done : 0.0;
    for (int64_t i = 0; i < nv; i++) {
        const size_t k = (size_t)i;
        rgb[i] = (((double)s[k] - (double)c[k]) + ((double)ds[k] - (double)dc[k])) * inv;
    }
    return PPPM_OK;
}

int pppm_image_save(void* ctx_raw, const char* path, double gamma) {
    CHECK_CTX(ctx);
    if (!ctx->rendering) return ctx->fail(PPPM_ERR_STATE, "no render in progress");
    if (!path) return ctx->fail(PPPM_ERR_INVALID, "no path");
    std::string ext(path);
    const size_t dot = ext.find_last_of('.');
    ext = dot == std::string::npos ? "" : ext.substr(dot);
    for (char& ch : ext) ch = (char)std::tolower((unsigned char)ch);
    const int fmt = (ext == ".png" || ext == ".bmp") ? 0 : ext == ".hdr" ? 1 : ext == ".pfm" ? 2 : -1;
    if (fmt < 0) return ctx->fail(PPPM_ERR_INVALID, "images are saved as .png, .bmp, .hdr or .pfm");
    (void)hipSetDevice(ctx->device);
    const int32_t w = ctx->cam.width, h = ctx->cam.height;
    const int64_t np = (int64_t)w * h;
    if (ctx->rp.method == PPPM_METHOD_HACHISUKA) {  // per-point radii: the FP64 image, encoded on the host
        std::vector<double> img(3 * (size_t)np);
        const int rc = pppm_image_read(ctx_raw, img.data(), (int64_t)img.size());
        if (rc != PPPM_OK) return rc;
        const int wr = pppm_image_write(path, img.data(), w, h, gamma);
        return wr == PPPM_OK ? PPPM_OK : ctx->fail(wr, std::string("cannot write ") + path);
    }
    const size_t bytes = fmt == 0 ? 3 * (size_t)np : 4 * (size_t)np;  // RGB8, RGBE or 3 x FP32 / 3 bytes
    const size_t out_bytes = fmt == 2 ? 12 * (size_t)np : bytes;
    TRY_HIP(ctx, hipStreamSynchronize(ctx->stream2));
    TRY_HIP(ctx, ctx->scratch.ensure(out_bytes));
    const double inv = ctx->passes_done > 0 ? 1.0 / (double)ctx->passes_done : 0.0;
    launch_image_encode(ctx->stream, ctx->img_sum.as<float>(), ctx->img_comp.as<float>(), ctx->img_dsum.as<float>(),
                        ctx->img_dcomp.as<float>(), np, inv, fmt, gamma > 0.0 ? 1.0 / gamma : 0.0,
                        ctx->scratch.as<uint8_t>(), ctx->scratch.as<float>());
    TRY_HIP(ctx, hipGetLastError());
    std::vector<unsigned char> host(out_bytes);
    TRY_HIP(ctx, hipMemcpyAsync(host.data(), ctx->scratch.p, out_bytes, hipMemcpyDeviceToHost, ctx->stream));
    TRY_HIP(ctx, hipStreamSynchronize(ctx->stream));
    int rc;
    if (fmt == 0) rc = ext == ".png" ? image_write_png8(path, host.data(), w, h) : image_write_bmp8(path, host.data(), w, h);
    else if (fmt == 1) rc = image_write_rgbe(path, host.data(), w, h);
    else rc = image_write_pfm(path, reinterpret_cast<const float*>(host.data()), w, h);
    if (rc != PPPM_OK) return ctx->fail(rc, std::string("cannot write ") + path);
    return PPPM_OK;
}

int pppm_last_photons(void* ctx_raw, float* pos, float* flux, float* normal, uint32_t* key,
                      int64_t cap, int64_t* n) {
    CHECK_CTX(ctx);
    if (!ctx->last_valid) return ctx->fail(PPPM_ERR_STATE, "no completed pass");
    if (n) *n = ctx->last_n;
    const int64_t m = ctx->last_n < cap ? ctx->last_n : cap;
    if (m <= 0) return PPPM_OK;
    (void)hipSetDevice(ctx->device);
    TRY_HIP(ctx, ctx->scratch.ensure(sizeof(float4) * 2 * (size_t)m));
    float4* a = ctx->scratch.as<float4>();
    float4* b = a + m;
    launch_permute2(ctx->stream, ctx->ph_rec.as<float4>(), ctx->ph_rec.as<float4>() + 1, ctx->last_perm1,
                    nullptr, m, a, b, 2);
    std::vector<float4> ha((size_t)m), hb((size_t)m);
    TRY_HIP(ctx, hipMemcpyAsync(ha.data(), a, sizeof(float4) * (size_t)m, hipMemcpyDeviceToHost, ctx->stream));
    TRY_HIP(ctx, hipMemcpyAsync(hb.data(), b, sizeof(float4) * (size_t)m, hipMemcpyDeviceToHost, ctx->stream));
    TRY_HIP(ctx, hipStreamSynchronize(ctx->stream));
    for (int64_t i = 0; i < m; i++) {
        if (pos) { pos[3 * i] = ha[i].x; pos[3 * i + 1] = ha[i].y; pos[3 * i + 2] = ha[i].z; }
        if (flux) { flux[3 * i] = hb[i].x; flux[3 * i + 1] = hb[i].y; flux[3 * i + 2] = hb[i].z; }
        if (key) std::memcpy(&key[i], &hb[i].w, 4);
        if (normal) {
            uint32_t e;
            std::memcpy(&e, &ha[i].w, 4);
            oct_decode_host(e, normal + 3 * i);
        }
    }
    return PPPM_OK;
}

int pppm_last_grid(void* ctx_raw, int32_t* sorted_idx, int32_t* cell_start, int32_t* cell_end,
                   int32_t hash_size, double bbox_min[3], double* hash_scale) {
    CHECK_CTX(ctx);
    if (!ctx->last_valid) return ctx->fail(PPPM_ERR_STATE, "no completed pass");
    if (hash_size != ctx->last_S) return ctx->fail(PPPM_ERR_INVALID, "hash_size must equal the pass's");
    (void)hipSetDevice(ctx->device);
    std::vector<uint32_t> off((size_t)hash_size + 1);
    const size_t n = (size_t)std::max<int64_t>(ctx->last_n, 0);
    std::vector<uint32_t> slot(sorted_idx ? n : 0), perm(sorted_idx ? n : 0);
    if (sorted_idx && n > 0) {
        TRY_HIP(ctx, hipMemcpyAsync(slot.data(), ctx->last_sorted, sizeof(uint32_t) * n, hipMemcpyDeviceToHost, ctx->stream));
        TRY_HIP(ctx, hipMemcpyAsync(perm.data(), ctx->last_perm1, sizeof(uint32_t) * n, hipMemcpyDeviceToHost, ctx->stream));
    }
    TRY_HIP(ctx, hipMemcpyAsync(off.data(), ctx->off.p, sizeof(uint32_t) * off.size(), hipMemcpyDeviceToHost, ctx->stream));
    TRY_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if (sorted_idx && n > 0) {
        // the grid sort carried record slots; a record's id is its origin
        // position, the inverse of the origin order slot = perm[position]
        std::vector<uint32_t> origin(n, UINT32_MAX);
        for (size_t p = 0; p < n; p++) {
            if (perm[p] >= n || origin[perm[p]] != UINT32_MAX)
                return ctx->fail(PPPM_ERR_STATE, "origin order is not a permutation");
            origin[perm[p]] = (uint32_t)p;
        }
        for (size_t k = 0; k < n; k++) {
            if (slot[k] >= n) return ctx->fail(PPPM_ERR_STATE, "grid value out of range");
            sorted_idx[k] = (int32_t)origin[slot[k]];
        }
    }
    if (sorted_idx && ctx->last_sub_bits > 0) {
        // the pass kept each bucket in cell-local Morton order for the gather;
        // the record ids of a bucket in ascending order are its insertion order
        for (int64_t h = 0; h < hash_size; h++)
            std::sort(sorted_idx + off[(size_t)h], sorted_idx + off[(size_t)h + 1]);
    }
    copy_ranges(off, hash_size, cell_start, cell_end);
    if (bbox_min)
        for (int d = 0; d < 3; d++) bbox_min[d] = ctx->grid_bmin[d];
    if (hash_scale) *hash_scale = ctx->last_scale;
    return PPPM_OK;
}

int pppm_last_hitpoints(void* ctx_raw, float* pos, float* normal, float* weight, int32_t* pixel,
                        int64_t cap, int64_t* n) {
    CHECK_CTX(ctx);
    if (!ctx->last_valid) return ctx->fail(PPPM_ERR_STATE, "no completed pass");
    (void)hipSetDevice(ctx->device);
    Counters hc;
    TRY_HIP(ctx, hipMemcpy(&hc, ctx->counters.p, sizeof(Counters), hipMemcpyDeviceToHost));
    const int64_t nh = (int64_t)hc.hitpoints;
    if (n) *n = nh;
    const int64_t m = nh < cap ? nh : cap;
    if (m <= 0) return PPPM_OK;
    std::vector<float4> a((size_t)m), b((size_t)m);
    // the compacted, Morton-ordered hit points the gather consumed
    TRY_HIP(ctx, hipMemcpy(a.data(), ctx->hp_pos_s.p, sizeof(float4) * (size_t)m, hipMemcpyDeviceToHost));
    TRY_HIP(ctx, hipMemcpy(b.data(), ctx->hp_w_s.p, sizeof(float4) * (size_t)m, hipMemcpyDeviceToHost));
    for (int64_t i = 0; i < m; i++) {
        if (pos) { pos[3 * i] = a[i].x; pos[3 * i + 1] = a[i].y; pos[3 * i + 2] = a[i].z; }
        if (normal) {
            uint32_t e;
            std::memcpy(&e, &a[i].w, 4);
            oct_decode_host(e, normal + 3 * i);
        }
        if (weight) { weight[3 * i] = b[i].x; weight[3 * i + 1] = b[i].y; weight[3 * i + 2] = b[i].z; }
        if (pixel) std::memcpy(&pixel[i], &b[i].w, 4);
    }
    return PPPM_OK;
}

int pppm_gather(void* ctx_raw, const float* ph_pos, const float* ph_flux, const float* ph_nrm,
                int64_t n_ph, const float* hit_pos, const float* hit_nrm, const float* hit_w,
                int64_t n_hit, double radius, int32_t hash_size, const double bbox_min[3],
                int32_t gather_mode, float* out_rgb, pppm_pass_stats* st) {
    CHECK_CTX(ctx);
    if (gather_mode < PPPM_GATHER_AUTO || gather_mode > PPPM_GATHER_BLOCKS)
        return ctx->fail(PPPM_ERR_INVALID, "unknown gather mode");
    if (gather_mode == PPPM_GATHER_AUTO) gather_mode = GATHER_AUTO_CHOICE;
    if (n_ph < 0 || n_hit < 0 || hash_size <= 0 || !(radius > 0.0) || !bbox_min ||
        (n_ph > 0 && (!ph_pos || !ph_flux || !ph_nrm)) ||
        (n_hit > 0 && (!hit_pos || !hit_nrm || !hit_w || !out_rgb)) || n_ph > 0x7FFFFFFFLL)
        return ctx->fail(PPPM_ERR_INVALID, "bad gather arguments");
    (void)hipSetDevice(ctx->device);
    hipStream_t s = ctx->stream;
    const int64_t np = n_ph > 0 ? n_ph : 1, nh = n_hit > 0 ? n_hit : 1;
    // staging: photons [3 x 3 floats] + hit points [3 x 3 floats] + outputs
    // (the lanes kernel's reads past the sorted positions p4s land in f4s and
    // what follows; GATHER_PAD records at the end keep them inside)
    size_t bytes = sizeof(float) * 9 * (size_t)(np + nh) + sizeof(float4) * 5 * (size_t)np +
                   sizeof(float4) * 2 * (size_t)nh + sizeof(float) * 3 * (size_t)nh + sizeof(float4) * GATHER_PAD;
    TRY_HIP(ctx, ctx->scratch.ensure(bytes));
    float* f = ctx->scratch.as<float>();
    float* dpp = f; float* dpf = dpp + 3 * np; float* dpn = dpf + 3 * np;
    float* dhp = dpn + 3 * np; float* dhn = dhp + 3 * nh; float* dhw = dhn + 3 * nh;
    float4* p4 = reinterpret_cast<float4*>(dhw + 3 * nh);
    float4* f4 = p4 + np;
    float4* p4s = f4 + np;
    float4* f4s = p4s + np;
    float4* c4s = f4s + np;
    float4* h4 = c4s + np;
    float4* w4 = h4 + nh;
    float* dout = reinterpret_cast<float*>(w4 + nh);
    if (n_ph > 0) {
        TRY_HIP(ctx, hipMemcpyAsync(dpp, ph_pos, sizeof(float) * 3 * (size_t)n_ph, hipMemcpyHostToDevice, s));
        TRY_HIP(ctx, hipMemcpyAsync(dpf, ph_flux, sizeof(float) * 3 * (size_t)n_ph, hipMemcpyHostToDevice, s));
        TRY_HIP(ctx, hipMemcpyAsync(dpn, ph_nrm, sizeof(float) * 3 * (size_t)n_ph, hipMemcpyHostToDevice, s));
    }
    if (n_hit > 0) {
        TRY_HIP(ctx, hipMemcpyAsync(dhp, hit_pos, sizeof(float) * 3 * (size_t)n_hit, hipMemcpyHostToDevice, s));
        TRY_HIP(ctx, hipMemcpyAsync(dhn, hit_nrm, sizeof(float) * 3 * (size_t)n_hit, hipMemcpyHostToDevice, s));
        TRY_HIP(ctx, hipMemcpyAsync(dhw, hit_w, sizeof(float) * 3 * (size_t)n_hit, hipMemcpyHostToDevice, s));
    }
    launch_pack_points(s, dpp, dpf, dpn, n_ph, p4, f4, 1);
    launch_pack_points(s, dhp, dhw, dhn, n_hit, h4, w4, 1);
    TRY_HIP(ctx, ensure_sort_buffers(ctx, n_ph, (uint32_t)hash_size));
    TRY_HIP(ctx, ctx->counters.ensure(sizeof(Counters)));
    Counters* dc = ctx->counters.as<Counters>();
    launch_zero_u32(s, dc, sizeof(Counters) / 4);
    GridParams g;
    for (int d = 0; d < 3; d++) g.bmin[d] = bbox_min[d];
    g.scale = 1.0 / (2.0 * radius);
    g.size = (uint32_t)hash_size;
    // the lanes kernel as in a pass: Morton order inside buckets, block boxes
    if (gather_compact(gather_mode))
        g.sub_bits = std::max(0, std::min(3, (32 - grid_bucket_bits(g.size, true)) / 3));
    uint32_t* sorted = build_grid(ctx, p4, nullptr, nullptr, n_ph, g, -1);
    launch_gather_layout(s, p4, f4, nullptr, sorted, n_ph, p4s, f4s, gather_compact(gather_mode) ? nullptr : c4s);
    if (g.sub_bits > 0) {
        TRY_HIP(ctx, ctx->boxes.ensure(sizeof(float4) * 2 * (size_t)gather_box_count(n_ph)));
        launch_gather_boxes(s, p4s, n_ph, ctx->boxes.as<float4>());
    }
    GatherArgs ga;
    ga.mode = gather_mode;
    ga.ph_a = p4s;
    ga.ph_b = f4s;
    ga.ph_c = c4s;
    ga.off = ctx->off.as<uint32_t>();
    if (n_hit > 0) {  // every kernel reads the hit points in Morton order, as in a pass
        TRY_HIP(ctx, ensure_hit_sort(ctx, n_hit));
        sort_hitpoints(ctx, ctx->stream, h4, w4, n_hit, g, radius);
        ga.hp_pos = ctx->hp_pos_s.as<float4>();
        ga.hp_w = ctx->hp_w_s.as<float4>();
    } else {
        ga.hp_pos = h4;
        ga.hp_w = w4;
    }
    TRY_HIP(ctx, hipEventRecord(ctx->ev[0], s));
    ga.hp_count_src = nullptr;
    ga.n_hit = n_hit;
    ga.g = g;
    ga.radius = radius;
    ga.img_sum = nullptr;
    ga.img_comp = nullptr;
    ga.out_rgb = dout;
    ga.counters = dc;
    ga.used = nullptr;
    ga.boxes = g.sub_bits > 0 ? ctx->boxes.as<float4>() : nullptr;
    launch_gather(s, ga);
    TRY_HIP(ctx, hipGetLastError());
    TRY_HIP(ctx, hipEventRecord(ctx->ev[1], s));
    if (n_hit > 0)
        TRY_HIP(ctx, hipMemcpyAsync(out_rgb, dout, sizeof(float) * 3 * (size_t)n_hit, hipMemcpyDeviceToHost, s));
    Counters hc;
    TRY_HIP(ctx, hipMemcpyAsync(&hc, dc, sizeof(Counters), hipMemcpyDeviceToHost, s));
    TRY_HIP(ctx, hipStreamSynchronize(s));
    if (st) {
        std::memset(st, 0, sizeof(*st));
        float ms = 0.0f;
        (void)hipEventElapsedTime(&ms, ctx->ev[0], ctx->ev[1]);
        st->gather_ms = ms;
        st->radius = radius;
        st->hash_size = hash_size;
        st->photons_stored = n_ph;
        st->hitpoints = n_hit;
        st->gather_buckets = (int64_t)hc.gather_buckets;
        st->gather_candidates = (int64_t)hc.gather_candidates;
        st->gather_inradius = (int64_t)hc.gather_inradius;
        st->gather_tested = (int64_t)hc.gather_tested;
        st->gather_blocks = (int64_t)hc.gather_blocks;
    }
    return PPPM_OK;
}

int pppm_stream_copy(void* ctx_raw, int64_t bytes, double* ms) {
    CHECK_CTX(ctx);
    if (bytes < 16) return ctx->fail(PPPM_ERR_INVALID, "copy at least 16 bytes");
    (void)hipSetDevice(ctx->device);
    const int64_t n = bytes / 16;
    void* a = nullptr;
    void* b = nullptr;
    TRY_HIP(ctx, hipMalloc(&a, (size_t)n * 16));
    hipError_t e = hipMalloc(&b, (size_t)n * 16);
    if (e != hipSuccess) {
        (void)hipFree(a);
        return ctx->fail(PPPM_ERR_HIP, std::string("hipMalloc: ") + hipGetErrorString(e));
    }
    launch_zero_u32(ctx->stream, a, 4 * n);
    (void)hipEventRecord(ctx->ev[0], ctx->stream);
    launch_stream_copy(ctx->stream, reinterpret_cast<const float4*>(a), reinterpret_cast<float4*>(b), n);
    (void)hipEventRecord(ctx->ev[1], ctx->stream);
    e = hipEventSynchronize(ctx->ev[1]);
    float t = 0.0f;
    (void)hipEventElapsedTime(&t, ctx->ev[0], ctx->ev[1]);
    (void)hipFree(a);
    (void)hipFree(b);
    if (e != hipSuccess) return ctx->fail(PPPM_ERR_HIP, std::string("copy: ") + hipGetErrorString(e));
    if (ms) *ms = t;
    return PPPM_OK;
}

int pppm_kahan_probe(void* ctx_raw, const float* values, int64_t n, int32_t lanes, float* sum, float* comp) {
    CHECK_CTX(ctx);
    if (!values || n <= 0 || lanes < 1 || lanes > 64 || !sum || !comp)
        return ctx->fail(PPPM_ERR_INVALID, "kahan probe: n > 0 values and 1..64 lanes");
    (void)hipSetDevice(ctx->device);
    void* buf = nullptr;
    TRY_HIP(ctx, hipMalloc(&buf, sizeof(float) * ((size_t)n + 128)));
    float* dv = static_cast<float*>(buf);
    float* ds = dv + n;
    float* dc = ds + 64;
    hipError_t e = hipMemcpyAsync(dv, values, sizeof(float) * (size_t)n, hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess) {
        launch_kahan_probe(ctx->stream, dv, n, lanes, ds, dc);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpyAsync(sum, ds, sizeof(float) * (size_t)lanes, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(comp, dc, sizeof(float) * (size_t)lanes, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    (void)hipFree(buf);
    if (e != hipSuccess) return ctx->fail(PPPM_ERR_HIP, std::string("kahan probe: ") + hipGetErrorString(e));
    return PPPM_OK;
}

int pppm_issue_probe(void* ctx_raw, double* valu_cycles, double* trans_cycles, double* clock_hz) {
    CHECK_CTX(ctx);
    (void)hipSetDevice(ctx->device);
    int clk_khz = 0;
    TRY_HIP(ctx, hipDeviceGetAttribute(&clk_khz, hipDeviceAttributeClockRate, ctx->device));
    const int blocks = (ctx->cu > 0 ? ctx->cu : 256) * 16;
    float* out = nullptr;
    TRY_HIP(ctx, hipMalloc(&out, sizeof(float) * 256 * (size_t)blocks));
    double cyc[2] = {0.0, 0.0};
    hipError_t e = hipSuccess;
    for (int which = 0; which < 2 && e == hipSuccess; which++) {
        float best = 1e30f;
        for (int rep = 0; rep < 3 && e == hipSuccess; rep++) {
            (void)hipEventRecord(ctx->ev[0], ctx->stream);
            launch_issue_probe(ctx->stream, out, blocks, which);
            (void)hipEventRecord(ctx->ev[1], ctx->stream);
            e = hipEventSynchronize(ctx->ev[1]);
            float ms = 0.0f;
            (void)hipEventElapsedTime(&ms, ctx->ev[0], ctx->ev[1]);
            if (ms > 0.0f) best = std::min(best, ms);
        }
        // cycles per wave instruction on one SIMD at the reported clock
        const double waves = (double)blocks * 4.0;
        const double rate = waves * (double)issue_probe_wave_instr(which) / (best * 1e-3);
        cyc[which] = (double)(ctx->cu > 0 ? ctx->cu : 256) * 4.0 * clk_khz * 1e3 / rate;
    }
    (void)hipFree(out);
    if (e != hipSuccess) return ctx->fail(PPPM_ERR_HIP, std::string("probe: ") + hipGetErrorString(e));
    if (valu_cycles) *valu_cycles = cyc[0];
    // the exp probe alternates v_exp_f32 and a full-rate v_fma_f32
    if (trans_cycles) *trans_cycles = 2.0 * cyc[1] - cyc[0];
    if (clock_hz) *clock_hz = clk_khz * 1e3;
    return PPPM_OK;
}

}  // extern "C"
