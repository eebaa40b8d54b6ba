// capi.hip -- the C ABI of libpppm_hip.so (declared in include/pppm.h).
//
// Owns the device state of one renderer: scene, 4-wide BVH, per-pass photon /
// hit-point buffers and the Kahan-compensated FP32 image.  The pass loop of
// ProgressivePhotonMappingProb::render (sources/ppm_probabilistic.cc:24-121)
// is split into pppm_render_begin / pppm_render_pass / pppm_image_read.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "internal.h"
#include "traverse.h"

namespace {

thread_local std::string g_create_error;

struct DBuf {
    void* p = nullptr;
    size_t n = 0;
    DBuf() = default;
    DBuf(const DBuf&) = delete;
    DBuf& operator=(const DBuf&) = delete;
    ~DBuf() { release(); }
    hipError_t ensure(size_t bytes) {
        if (bytes == 0) bytes = 16;
        if (bytes <= n) return hipSuccess;
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
        hipError_t e = hipMalloc(&p, bytes);
        if (e == hipSuccess) n = bytes;
        return e;
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    template <class T>
    T* as() const {
        return reinterpret_cast<T*>(p);
    }
};

struct Ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    int cu = 0;
    std::string name;
    size_t hbm = 0;
    // scene
    bool has_scene = false;
    int64_t n_tris = 0;
    DBuf tri9, tri_mat, mats, lights;
    std::vector<pppm_bsdf> h_bsdf;
    std::vector<pppm_light> h_lights;
    double bmin[3] = {0, 0, 0}, bmax[3] = {0, 0, 0};
    // BVH
    bool has_bvh = false;
    Node4* nodes = nullptr;
    float4* tris4 = nullptr;
    int32_t root = EMPTY_REF;
    pppm_bvh_info binfo{};
    // camera
    bool has_cam = false;
    pppm_camera cam{};
    // render
    bool rendering = false;
    pppm_render_params rp{};
    int32_t passes_done = 0;
    std::vector<double> radii;
    int64_t cap = 0;
    double grid_bmin[3] = {0, 0, 0};
    // ph_rec: photons as the photon pass appended them, slot j's position at
    // ph_rec[2 j] and its flux at ph_rec[2 j + 1] (one record, one 32-byte read);
    // ph_a/b/c: the gather layout in bucket order
    // img_sum / img_comp: the gather (and the dipole sum); img_dsum / img_dcomp:
    // emission seen directly by eye paths (its own Kahan pair, so the eye pass
    // never writes the gather's accumulator)
    DBuf ph_rec, ph_a, ph_b, ph_c, u[5], hist, off, hp_pos, hp_w, img_sum, img_comp, img_dsum, img_dcomp,
        counters, scratch;
    // hit points in Morton order of their half cells (tiled gather)
    DBuf hp_pos_s, hp_w_s, hk[4], hhist;
    // the photon records' origin order without a sort: rank of each record
    // among its photon's records, records per photon, their exclusive scan
    DBuf rec_rank, rec_count, rec_start, rec_tiles;
    // last pass
    bool last_valid = false;
    int64_t last_n = 0;
    int32_t last_S = 0;
    double last_scale = 0.0;
    uint32_t* last_perm1 = nullptr;
    uint32_t* last_sorted = nullptr;
    int32_t last_sub_bits = 0;  // the last pass ordered each bucket by a cell-local Morton code
    int32_t count_traversal = 0;  // passes run the node / triangle counting kernels
    hipEvent_t ev[13] = {};
    hipStream_t stream2 = nullptr;  // the eye pass runs here, concurrent with the photon pass
    hipStream_t stream3 = nullptr;  // the gather of pipelined passes (pppm_render_passes)
    hipStream_t stream4 = nullptr;  // the eye passes of odd pipelined passes
    hipEvent_t pev[6][11] = {};     // events of pipelined passes (a ring over passes)
    // the second buffer set of pipelined passes: what a pass's gather reads
    DBuf ph_a_alt, ph_b_alt, ph_c_alt, off_alt, hp_pos_s_alt, hp_w_s_alt, counters_alt, pass_counters, boxes_alt;
    DBuf ph_a_alt2, ph_b_alt2, ph_c_alt2, off_alt2, hp_pos_s_alt2, hp_w_s_alt2, counters_alt2, boxes_alt2;
    DBuf hp_pos_alt, hp_w_alt, hp_pos_alt2, hp_w_alt2;  // eye-pass pixel slots per set
    DBuf hk_alt[4], hhist_alt;                          // hit point sort of the second eye stream
    // block boxes of the lanes gather over ph_a (GATHER_BLOCK records each)
    DBuf boxes;
    // subsurface scattering (BSSRDF materials)
    DBuf dips, sss_pts, sss_ptw, sss_E, irr_pos, irr_flux, sss_pos, sss_w;
    std::vector<int32_t> h_sss_index;  // material -> dipole index (-1: none)
    int64_t sss_points_req = 0;        // 0: the default count
    int64_t sss_n = 0;                 // surface points of the current render
    double sss_area = 0.0;             // area per point
    bool sss_present = false;
    int32_t n_dips = 0;
    int64_t last_n_irr = 0;
    // device pass state for passes enqueued without host synchronisation
    DBuf pdev, radii_dev;
    // learned emission proposal
    struct Proposal {
        bool enabled = false;  // records + retraining
        bool active = false;   // the photon kernel samples the tables
        pppm_proposal_params pp{};
        int32_t side = 0;      // anchors per light edge of the current tables
        float lambda = 1.0f;
        int64_t ring_cap = 0, ring_head = 0, ring_valid = 0;
        int32_t adam_t = 0;
        uint32_t draws = 0;
        double last_loss = 0.0;
        bool mlp_ready = false;
    } prop;
    DBuf mlp_params, mlp_m, mlp_v, mlp_wb, mlp_part, mlp_act, mlp_idx, mlp_wsum, mlp_loss, mlp_stage,
        mlp_grad, prop_ring, prop_prob, prop_cdf, prop_uv, prop_flags, prop_used, prop_segc, prop_scan, prop_bsums;

    int fail(int code, const std::string& msg) {
        err = msg;
        return code;
    }
    SceneDev scene_dev() const {
        SceneDev s;
        s.nodes = nodes;
        s.tris = tris4;
        s.tri_mat = tri_mat.as<int32_t>();
        s.mats = mats.as<Mat>();
        s.lights = lights.as<LightDev>();
        s.n_lights = (int32_t)h_lights.size();
        s.root = root;
        s.n_nodes = (int32_t)binfo.nodes;
        s.n_tris = n_tris;
        return s;
    }
};

#define TRY_HIP(ctx, x)                                                                    \
    do {                                                                                   \
        hipError_t e_ = (x);                                                               \
        if (e_ != hipSuccess)                                                              \
            return (ctx)->fail(PPPM_ERR_HIP, std::string(#x) + ": " + hipGetErrorString(e_)); \
    } while (0)

#define CHECK_CTX(c)                                   \
    Ctx* c = reinterpret_cast<Ctx*>(c##_raw);          \
    if (!c) return PPPM_ERR_INVALID;                   \
    c->err.clear()

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
int render_sub_bits(const Ctx* c, uint32_t size, bool sentinel) {
    if (c->rp.gather_mode != GATHER_LANES) return 0;
    return std::max(0, std::min(3, (32 - grid_bucket_bits(size, sentinel)) / 3));
}

uint32_t* build_grid(Ctx* c, const float4* pos4, const float* pos3, const uint32_t* perm, int64_t n,
                     const GridParams& g, int keep, const unsigned long long* count = nullptr,
                     uint32_t* off_out = nullptr, bool keys_ready = false) {
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
uint32_t* origin_order(Ctx* c, hipStream_t s, int64_t n, const GridParams& g, const unsigned long long* count) {
    const int bits = grid_bucket_bits(g.size, count != nullptr) + 3 * g.sub_bits;
    const uint32_t sentinel = bits >= 32 ? 0xFFFFFFFFu : (uint32_t)((1ull << bits) - 1);
    launch_origin_starts(s, c->rec_count.as<uint32_t>(), c->rp.photons, c->rec_tiles.as<uint32_t>(),
                         c->rec_start.as<uint32_t>());
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
    c->sss_present = true;
    return hipSuccess;
}

// ---------------------------------------------------------------------------
// Learned emission proposal: host side
// ---------------------------------------------------------------------------
uint64_t splitmix64(uint64_t* st) {
    uint64_t z = (*st += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// He-uniform hidden layers; zero biases and output layer (uniform cells at first)
std::vector<float> mlp_init_params(uint64_t seed) {
    std::vector<float> w(MLP_NPARAM, 0.0f);
    uint64_t st = seed ^ 0x5851F42D4C957F2Dull;
    auto fill = [&](int off, int count, int fan_in) {
        const double a = std::sqrt(6.0 / fan_in);
        for (int i = 0; i < count; i++) {
            const double u = (double)(splitmix64(&st) >> 11) * (1.0 / 9007199254740992.0);
            w[(size_t)(off + i)] = (float)((2.0 * u - 1.0) * a);
        }
    };
    fill(OW1, MLP_HID * MLP_IN, MLP_IN);
    fill(OW2, MLP_HID * MLP_HID, MLP_HID);
    fill(OW3, MLP_HID * MLP_HID, MLP_HID);
    return w;
}

// Master parameters from the host; Adam state cleared; BF16 mirror rebuilt
hipError_t mlp_upload(Ctx* c, const float* params) {
    hipError_t e;
    for (DBuf* b : {&c->mlp_params, &c->mlp_m, &c->mlp_v, &c->mlp_grad})
        if ((e = b->ensure(sizeof(float) * MLP_NPARAM)) != hipSuccess) return e;
    if ((e = c->mlp_wb.ensure(sizeof(uint16_t) * MLP_NMIRROR)) != hipSuccess) return e;
    if ((e = c->mlp_part.ensure(sizeof(float) * MLP_SPLITS * (size_t)MLP_NPARAM)) != hipSuccess) return e;
    for (DBuf* b : {&c->mlp_wsum, &c->mlp_loss})
        if ((e = b->ensure(16)) != hipSuccess) return e;
    if ((e = hipMemcpyAsync(c->mlp_params.p, params, sizeof(float) * MLP_NPARAM, hipMemcpyHostToDevice,
                            c->stream)) != hipSuccess)
        return e;
    launch_zero_u32(c->stream, c->mlp_m.p, MLP_NPARAM);
    launch_zero_u32(c->stream, c->mlp_v.p, MLP_NPARAM);
    launch_mlp_mirror(c->stream, c->mlp_params.as<float>(), c->mlp_wb.as<uint16_t>());
    c->prop.adam_t = 0;
    c->prop.mlp_ready = true;
    return hipGetLastError();
}

hipError_t ensure_mlp(Ctx* c) {
    if (c->prop.mlp_ready) return hipSuccess;
    std::vector<float> w = mlp_init_params(0x7A75);
    hipError_t e = mlp_upload(c, w.data());
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    return e;
}

hipError_t ensure_train(Ctx* c, int64_t B) {
    hipError_t e = c->mlp_act.ensure(sizeof(uint16_t) * ACT_ROWS * (size_t)B);
    if (e == hipSuccess) e = c->mlp_idx.ensure(sizeof(uint32_t) * (size_t)B);
    return e;
}

// One training step on B records: drawn uniformly from rec[0, n_valid) or the
// first B (identity).  Adam applies when lr > 0; the loss is accumulated.
void mlp_step(Ctx* c, const float4* rec, int64_t n_valid, int64_t B, bool identity, double lr) {
    hipStream_t s = c->stream;
    launch_mlp_prep(s, rec, n_valid, B, c->prop.pp.seed, c->prop.draws, identity ? 1 : 0, c->mlp_idx.as<uint32_t>(),
                    c->mlp_wsum.as<float>());
    MlpArgs m{};
    m.params = c->mlp_params.as<float>();
    m.wb = c->mlp_wb.as<uint16_t>();
    m.rec = rec;
    m.idx = c->mlp_idx.as<uint32_t>();
    m.inv_wsum = c->mlp_wsum.as<float>();
    m.n = B;
    m.act = c->mlp_act.as<uint16_t>();
    m.loss = c->mlp_loss.as<float>();
    launch_mlp(s, m, true);
    launch_mlp_wgrad(s, c->mlp_act.as<uint16_t>(), B, c->mlp_part.as<float>());
    AdamArgs a{};
    a.params = c->mlp_params.as<float>();
    a.m = c->mlp_m.as<float>();
    a.v = c->mlp_v.as<float>();
    a.part = c->mlp_part.as<float>();
    a.part_stride = MLP_NPARAM;
    a.wb = c->mlp_wb.as<uint16_t>();
    a.grad_out = c->mlp_grad.as<float>();
    a.apply = lr > 0.0 ? 1 : 0;
    const int t = c->prop.adam_t > 0 ? c->prop.adam_t : 1;
    a.lr = (float)lr;
    a.b1 = 0.9f;
    a.b2 = 0.999f;
    a.eps = 1.0e-8f;
    a.c1 = (float)(1.0 / (1.0 - std::pow(0.9, t)));
    a.c2 = (float)(1.0 / (1.0 - std::pow(0.999, t)));
    launch_mlp_adam(s, a);
}

// Softmax of the MLP at the anchor centres -> per-anchor CDF tables
hipError_t proposal_tables_from_mlp(Ctx* c, int32_t side) {
    const int32_t A = side * side;
    std::vector<float> uv(2 * (size_t)A);
    for (int i = 0; i < side; i++)
        for (int j = 0; j < side; j++) {
            uv[2 * (size_t)(i * side + j)] = (float)((i + 0.5) / side);
            uv[2 * (size_t)(i * side + j) + 1] = (float)((j + 0.5) / side);
        }
    hipError_t e;
    if ((e = c->prop_uv.ensure(sizeof(float) * uv.size())) != hipSuccess) return e;
    if ((e = c->prop_prob.ensure(sizeof(float) * PROP_CELLS * (size_t)A)) != hipSuccess) return e;
    if ((e = c->prop_cdf.ensure(sizeof(float) * PROP_CELLS * (size_t)A)) != hipSuccess) return e;
    if ((e = hipMemcpyAsync(c->prop_uv.p, uv.data(), sizeof(float) * uv.size(), hipMemcpyHostToDevice, c->stream)) !=
        hipSuccess)
        return e;
    MlpArgs m{};
    m.params = c->mlp_params.as<float>();
    m.wb = c->mlp_wb.as<uint16_t>();
    m.uv = c->prop_uv.as<float2>();
    m.n = A;
    m.prob = c->prop_prob.as<float>();
    launch_mlp(c->stream, m, false);
    launch_prop_table(c->stream, c->prop_prob.as<float>(), A, c->prop_cdf.as<float>());
    if ((e = hipStreamSynchronize(c->stream)) != hipSuccess) return e;
    c->prop.side = side;
    return hipGetLastError();
}

// Retrain on the ring, then rebuild the tables; returns the last step's loss
hipError_t proposal_retrain(Ctx* c, double* loss) {
    const pppm_proposal_params& pp = c->prop.pp;
    hipError_t e = ensure_train(c, pp.batch);
    if (e != hipSuccess) return e;
    for (int k = 0; k < pp.train_steps; k++) {
        if (k == pp.train_steps - 1) launch_zero_u32(c->stream, c->mlp_loss.p, 1);
        c->prop.draws++;
        c->prop.adam_t++;
        mlp_step(c, c->prop_ring.as<float4>(), c->prop.ring_valid, pp.batch, false, pp.learning_rate);
    }
    float l = 0.0f;
    if ((e = hipMemcpyAsync(&l, c->mlp_loss.p, sizeof(float), hipMemcpyDeviceToHost, c->stream)) != hipSuccess)
        return e;
    if ((e = proposal_tables_from_mlp(c, pp.anchors)) != hipSuccess) return e;
    *loss = l;
    c->prop.lambda = (float)pp.lambda;
    c->prop.active = true;
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

// The hit points among the cap slots of (hp, hw), compacted into hp_pos_s /
// hp_w_s in Morton order of their half cells (edge r); a stable sort, so the
// order is a function of the slot order alone.
void sort_hitpoints(hipStream_t s, DBuf* hk, DBuf& hhist, const float4* hp, const float4* hw, int64_t cap,
                    const GridParams& g, double radius, float4* out_pos, float4* out_w) {
    uint32_t* k[4];
    for (int i = 0; i < 4; i++) k[i] = hk[i].as<uint32_t>();
    launch_hit_keys(s, hp, hw, cap, g, radius, k[0], k[1]);
    bool alt = radix_sort_pairs(s, k[0], k[1], k[2], k[3], hhist.as<uint32_t>(), cap, 32);
    launch_permute2(s, hp, hw, nullptr, alt ? k[3] : k[1], cap, out_pos, out_w);
}

void sort_hitpoints(Ctx* c, hipStream_t s, const float4* hp, const float4* hw, int64_t cap, const GridParams& g,
                    double radius) {
    sort_hitpoints(s, c->hk, c->hhist, hp, hw, cap, g, radius, c->hp_pos_s.as<float4>(), c->hp_w_s.as<float4>());
}

void copy_ranges(const std::vector<uint32_t>& off, int32_t S, int32_t* cs, int32_t* ce) {
    for (int32_t h = 0; h < S; h++) {
        if (cs) cs[h] = (int32_t)off[h];
        if (ce) ce[h] = (int32_t)off[h + 1];
    }
}

// Eye-pass arguments of the current render (camera basis in FP64, then FP32)
EyePassArgs eye_args(Ctx* ctx, uint32_t pass) {
    EyePassArgs ea;
    ea.sc = ctx->scene_dev();
    ea.cam = ctx->cam;
    const pppm_camera& c = ctx->cam;
    double D[3] = {c.direction[0], c.direction[1], c.direction[2]};
    double dn = std::sqrt(D[0] * D[0] + D[1] * D[1] + D[2] * D[2]);
    for (double& v : D) v /= dn;
    double up[3] = {c.up[0], c.up[1], c.up[2]};
    double U[3], V[3];
    cross3(D, up, U);
    double un = std::sqrt(U[0] * U[0] + U[1] * U[1] + U[2] * U[2]);
    for (double& v : U) v /= un;
    cross3(U, D, V);
    double th = std::tan(c.fov_y_deg * 0.5 * M_PI / 180.0);
    double aspect = (double)c.width / (double)c.height;
    for (int k = 0; k < 3; k++) {
        ea.cam_basis[k] = (float)c.center[k];
        ea.cam_basis[3 + k] = (float)D[k];
        ea.cam_basis[6 + k] = (float)(U[k] * th * aspect);
        ea.cam_basis[9 + k] = (float)(V[k] * th);
    }
    const int64_t npix = (int64_t)c.width * c.height;
    ea.seed = ctx->rp.seed;
    ea.pass = pass;
    ea.scene_id = ctx->rp.scene_id;
    ea.max_bounces = ctx->rp.max_bounces;
    ea.ray_eps = (float)ctx->rp.ray_eps;
    ea.hp_pos = ctx->hp_pos.as<float4>();
    ea.hp_w = ctx->hp_w.as<float4>();
    ea.img_sum = ctx->img_dsum.as<float>();
    ea.img_comp = ctx->img_dcomp.as<float>();
    ea.counters = ctx->counters.as<Counters>();
    ea.pd = nullptr;
    ea.sss_pos = ctx->sss_present ? ctx->sss_pos.as<float4>() : nullptr;
    ea.sss_w = ctx->sss_present ? ctx->sss_w.as<float4>() : nullptr;
    ea.sss_cap = ctx->sss_present ? npix : 0;
    ea.count_traversal = ctx->count_traversal;
    return ea;
}

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
    if (n < 0 || (n > 0 && (!tri_xyz || !bsdf_id)) || nb <= 0 || !table || nl <= 0 || !lights)
        return ctx->fail(PPPM_ERR_INVALID, "scene needs triangles, bsdf ids, a bsdf table and >= 1 light");
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

int pppm_bvh_build(void* ctx_raw, pppm_bvh_info* info) {
    CHECK_CTX(ctx);
    if (!ctx->has_scene) return ctx->fail(PPPM_ERR_STATE, "no scene uploaded");
    (void)hipSetDevice(ctx->device);
    (void)hipFree(ctx->nodes);
    (void)hipFree(ctx->tris4);
    ctx->nodes = nullptr;
    ctx->tris4 = nullptr;
    ctx->has_bvh = false;
    TRY_HIP(ctx, hipEventRecord(ctx->ev[0], ctx->stream));
    BvhBuild b;
    std::string e = bvh_build(ctx->stream, ctx->tri9.as<float>(), ctx->n_tris, &b);
    if (!e.empty()) return ctx->fail(PPPM_ERR_HIP, "BVH build: " + e);
    TRY_HIP(ctx, hipEventRecord(ctx->ev[1], ctx->stream));
    TRY_HIP(ctx, hipEventSynchronize(ctx->ev[1]));
    float ms = 0.0f;
    (void)hipEventElapsedTime(&ms, ctx->ev[0], ctx->ev[1]);
    ctx->nodes = b.nodes;
    ctx->tris4 = b.tris4;
    ctx->root = b.root;
    // the traversal stack holds at most 3 entries per level below the root
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

int pppm_render_begin(void* ctx_raw, const pppm_render_params* p) {
    CHECK_CTX(ctx);
    if (!ctx->has_bvh) return ctx->fail(PPPM_ERR_STATE, "no BVH built");
    if (!ctx->has_cam) return ctx->fail(PPPM_ERR_STATE, "no camera set");
    if (!p || p->photons <= 0 || p->passes <= 0 || !(p->alpha > 0.0 && p->alpha < 1.0) ||
        !(p->radius1 > 0.0) || p->max_bounces <= 0 || p->max_bounces > 64 || p->hash_size < 0 ||
        (p->gather_mode < PPPM_GATHER_AUTO || p->gather_mode > PPPM_GATHER_WAVE))
        return ctx->fail(PPPM_ERR_INVALID, "bad render parameters");
    if (p->photons > (1LL << 26)) return ctx->fail(PPPM_ERR_INVALID, "at most 2^26 photons per pass");
    (void)hipSetDevice(ctx->device);
    ctx->rp = *p;
    if (ctx->rp.gather_mode == PPPM_GATHER_AUTO) ctx->rp.gather_mode = GATHER_AUTO_CHOICE;
    ctx->radii.assign((size_t)p->passes, 0.0);
    pppm_radius_schedule(p->radius1, p->alpha, p->passes, 0, ctx->radii.data());
    double factor = p->capacity_factor > 0.0 ? p->capacity_factor : 8.0;
    ctx->cap = (int64_t)std::ceil((double)p->photons * factor);
    if (ctx->cap > 0x7FFFFFFFLL) ctx->cap = 0x7FFFFFFFLL;
    for (int d = 0; d < 3; d++) ctx->grid_bmin[d] = ctx->bmin[d] - p->radius1;
    // lights: selection by power, photon flux = P_total / N (FP64 on the host)
    int nl = (int)ctx->h_lights.size();
    std::vector<double> power(nl), area(nl);
    double total = 0.0;
    for (int l = 0; l < nl; l++) {
        const pppm_light& L = ctx->h_lights[l];
        double eu[3] = {L.edge_u[0], L.edge_u[1], L.edge_u[2]}, ev[3] = {L.edge_v[0], L.edge_v[1], L.edge_v[2]};
        double cr[3];
        cross3(eu, ev, cr);
        area[l] = std::sqrt(cr[0] * cr[0] + cr[1] * cr[1] + cr[2] * cr[2]);
        double lum = ((double)L.emission[0] + (double)L.emission[1] + (double)L.emission[2]) / 3.0;
        power[l] = M_PI * area[l] * lum;
        total += power[l];
    }
    if (!(total > 0.0)) return ctx->fail(PPPM_ERR_INVALID, "lights emit no power");
    std::vector<LightDev> ld(nl);
    double cum = 0.0;
    for (int l = 0; l < nl; l++) {
        const pppm_light& L = ctx->h_lights[l];
        double eu[3] = {L.edge_u[0], L.edge_u[1], L.edge_u[2]}, ev[3] = {L.edge_v[0], L.edge_v[1], L.edge_v[2]};
        double cr[3];
        cross3(eu, ev, cr);
        double inv = 1.0 / (area[l] + 1e-300);
        LightDev& o = ld[l];
        o.corner = mk(L.corner[0], L.corner[1], L.corner[2]);
        o.eu = mk(L.edge_u[0], L.edge_u[1], L.edge_u[2]);
        o.ev = mk(L.edge_v[0], L.edge_v[1], L.edge_v[2]);
        o.normal = mk((float)(cr[0] * inv), (float)(cr[1] * inv), (float)(cr[2] * inv));
        double pl = power[l] / total;
        double k = M_PI * area[l] / (pl * (double)p->photons);
        o.flux = mk((float)(L.emission[0] * k), (float)(L.emission[1] * k), (float)(L.emission[2] * k));
        cum += power[l];
        o.cdf_hi = (l == nl - 1) ? 2.0f : (float)(cum / total);
    }
    TRY_HIP(ctx, hipMemcpy(ctx->lights.p, ld.data(), sizeof(LightDev) * (size_t)nl, hipMemcpyHostToDevice));
    const int64_t npix = (int64_t)ctx->cam.width * ctx->cam.height;
    const int64_t cap = ctx->cap;
    TRY_HIP(ctx, ctx->ph_rec.ensure(sizeof(float4) * 2 * (size_t)cap));
    TRY_HIP(ctx, ctx->rec_rank.ensure((size_t)cap));
    TRY_HIP(ctx, ctx->rec_count.ensure(sizeof(uint32_t) * (size_t)p->photons));
    TRY_HIP(ctx, ctx->rec_start.ensure(sizeof(uint32_t) * (size_t)p->photons));
    TRY_HIP(ctx, ctx->rec_tiles.ensure(sizeof(uint32_t) * (size_t)origin_tiles(p->photons)));
    TRY_HIP(ctx, ctx->ph_a.ensure(sizeof(float4) * (size_t)(cap + GATHER_PAD)));
    TRY_HIP(ctx, ctx->ph_b.ensure(sizeof(float4) * (size_t)cap));
    TRY_HIP(ctx, ctx->ph_c.ensure(sizeof(float4) * (size_t)cap));
    TRY_HIP(ctx, ctx->boxes.ensure(sizeof(float4) * 2 * (size_t)gather_box_count(cap)));
    uint32_t Smax = p->hash_size > 0 ? (uint32_t)p->hash_size : (uint32_t)cap;
    TRY_HIP(ctx, ensure_sort_buffers(ctx, cap, Smax));
    TRY_HIP(ctx, ctx->hp_pos.ensure(sizeof(float4) * (size_t)npix));
    TRY_HIP(ctx, ctx->hp_w.ensure(sizeof(float4) * (size_t)npix));
    TRY_HIP(ctx, ensure_hit_sort(ctx, npix));
    TRY_HIP(ctx, ctx->img_sum.ensure(sizeof(float) * 3 * (size_t)npix));
    TRY_HIP(ctx, ctx->img_comp.ensure(sizeof(float) * 3 * (size_t)npix));
    TRY_HIP(ctx, ctx->img_dsum.ensure(sizeof(float) * 3 * (size_t)npix));
    TRY_HIP(ctx, ctx->img_dcomp.ensure(sizeof(float) * 3 * (size_t)npix));
    TRY_HIP(ctx, ctx->counters.ensure(sizeof(Counters)));
    launch_zero_u32(ctx->stream, ctx->img_sum.p, 3 * npix);
    launch_zero_u32(ctx->stream, ctx->img_comp.p, 3 * npix);
    launch_zero_u32(ctx->stream, ctx->img_dsum.p, 3 * npix);
    launch_zero_u32(ctx->stream, ctx->img_dcomp.p, 3 * npix);
    TRY_HIP(ctx, hipGetLastError());
    TRY_HIP(ctx, hipStreamSynchronize(ctx->stream));
    TRY_HIP(ctx, subsurface_setup(ctx, nullptr));
    if (ctx->sss_present) {
        TRY_HIP(ctx, ctx->irr_pos.ensure(sizeof(float4) * (size_t)cap));
        TRY_HIP(ctx, ctx->irr_flux.ensure(sizeof(float4) * (size_t)cap));
        TRY_HIP(ctx, ctx->sss_pos.ensure(sizeof(float4) * (size_t)npix));
        TRY_HIP(ctx, ctx->sss_w.ensure(sizeof(float4) * (size_t)npix));
    }
    ctx->passes_done = 0;
    ctx->rendering = true;
    ctx->last_valid = false;
    ctx->prop.enabled = false;
    ctx->prop.active = false;
    return PPPM_OK;
}

int pppm_render_pass(void* ctx_raw, pppm_pass_stats* st) {
    CHECK_CTX(ctx);
    if (!ctx->rendering) return ctx->fail(PPPM_ERR_STATE, "pppm_render_begin not called");
    if (ctx->passes_done >= ctx->rp.passes) return ctx->fail(PPPM_ERR_STATE, "all passes done");
    (void)hipSetDevice(ctx->device);
    hipStream_t s = ctx->stream;
    const int32_t pass = ctx->passes_done + 1;
    const double radius = ctx->radii[(size_t)pass - 1];
    const int64_t npix = (int64_t)ctx->cam.width * ctx->cam.height;
    const int eye_grid = ctx->cu > 0 ? ctx->cu * EYE_BLOCKS_PER_CU : 2048;
    const int photon_grid = ctx->cu > 0 ? ctx->cu * PHOTON_BLOCKS_PER_CU : 2048;
    Counters* dc = ctx->counters.as<Counters>();
    ctx->last_valid = false;

    launch_zero_u32(s, dc, sizeof(Counters) / 4);
    TRY_HIP(ctx, hipEventRecord(ctx->ev[0], s));
    // the eye pass and the hit point sort need no photons: they run on the
    // second stream, concurrent with the photon pass, so each kernel's tail of
    // long paths overlaps the other's work
    hipStream_t s2 = ctx->stream2;
    TRY_HIP(ctx, hipStreamWaitEvent(s2, ctx->ev[0], 0));
    TRY_HIP(ctx, hipEventRecord(ctx->ev[11], s2));
    {
        EyePassArgs ea = eye_args(ctx, (uint32_t)pass);
        launch_eye_pass(s2, ea, eye_grid);
        TRY_HIP(ctx, hipGetLastError());
        TRY_HIP(ctx, hipEventRecord(ctx->ev[4], s2));
        GridParams gh;
        for (int d = 0; d < 3; d++) gh.bmin[d] = ctx->grid_bmin[d];
        gh.scale = 1.0 / (2.0 * radius);
        gh.size = 1;
        sort_hitpoints(ctx, s2, ctx->hp_pos.as<float4>(), ctx->hp_w.as<float4>(), npix, gh, radius);
        TRY_HIP(ctx, hipGetLastError());
        TRY_HIP(ctx, hipEventRecord(ctx->ev[5], s2));
    }
    PhotonPassArgs pa;
    pa.sc = ctx->scene_dev();
    pa.n_photons = ctx->rp.photons;
    pa.cap = ctx->cap;
    pa.seed = ctx->rp.seed;
    pa.pass = (uint32_t)pass;
    pa.scene_id = ctx->rp.scene_id;
    pa.max_bounces = ctx->rp.max_bounces;
    pa.ray_eps = (float)ctx->rp.ray_eps;
    pa.ph_pos = ctx->ph_rec.as<float4>();
    pa.ph_flux = ctx->ph_rec.as<float4>() + 1;
    pa.rec_stride = 2;
    pa.counters = dc;
    const bool prop_active = ctx->prop.active;
    pa.prop.cdf = prop_active ? ctx->prop_cdf.as<float>() : nullptr;
    pa.prop.lambda = prop_active ? ctx->prop.lambda : 1.0f;
    pa.prop.side = prop_active ? ctx->prop.side : 0;
    pa.irr_pos = ctx->sss_present ? ctx->irr_pos.as<float4>() : nullptr;
    pa.irr_flux = ctx->sss_present ? ctx->irr_flux.as<float4>() : nullptr;
    pa.irr_cap = ctx->sss_present ? ctx->cap : 0;
    pa.count_traversal = ctx->count_traversal;
    pa.rec_rank = ctx->rec_rank.as<uint8_t>();
    pa.rec_count = ctx->rec_count.as<uint32_t>();
    launch_photon_pass(s, pa, photon_grid);
    TRY_HIP(ctx, hipGetLastError());
    TRY_HIP(ctx, hipEventRecord(ctx->ev[1], s));
    Counters hc;
    TRY_HIP(ctx, hipMemcpyAsync(&hc, dc, sizeof(Counters), hipMemcpyDeviceToHost, s));
    TRY_HIP(ctx, hipStreamSynchronize(s));
    const int64_t n = (int64_t)hc.photons_stored;
    if (st) {
        std::memset(st, 0, sizeof(*st));
        st->pass_index = pass;
        st->radius = radius;
        st->photons_emitted = ctx->rp.photons;
        st->photons_stored = n;
    }
    if (n > ctx->cap)
        return ctx->fail(PPPM_ERR_CAPACITY, "photon buffer holds " + std::to_string(ctx->cap) +
                                                " records, the pass stored " + std::to_string(n));
    const int64_t n_irr = (int64_t)hc.irr_stored;
    if (n_irr > ctx->cap)
        return ctx->fail(PPPM_ERR_CAPACITY, "irradiance photon buffer holds " + std::to_string(ctx->cap) +
                                                " records, the pass stored " + std::to_string(n_irr));
    TRY_HIP(ctx, hipEventRecord(ctx->ev[8], s));
    if (ctx->sss_present) {
        // irradiance at the surface points: the fixed-radius gather over the
        // photons stored on BSSRDF surfaces, with weight 1 / pi (the E_j of
        // sources/subsurface_integrator.cc:329-365)
        TRY_HIP(ctx, subsurface_irradiance(ctx, n_irr, radius));
    }
    TRY_HIP(ctx, hipEventRecord(ctx->ev[9], s));
    // hash grid over the photons in their deterministic order, (photon index,
    // bounce): the order and the grid keys in one pass, then the grid sort
    uint32_t S = ctx->rp.hash_size > 0 ? (uint32_t)ctx->rp.hash_size : (uint32_t)(n > 0 ? n : 1);
    GridParams g;
    for (int d = 0; d < 3; d++) g.bmin[d] = ctx->grid_bmin[d];
    g.scale = 1.0 / (2.0 * radius);
    g.size = S;
    g.sub_bits = render_sub_bits(ctx, S, false);
    TRY_HIP(ctx, ctx->off.ensure(sizeof(uint32_t) * ((size_t)S + 1)));
    const int keep = 4;
    uint32_t* perm1 = origin_order(ctx, s, n, g, nullptr);
    TRY_HIP(ctx, hipEventRecord(ctx->ev[2], s));
    uint32_t* sorted = build_grid(ctx, nullptr, nullptr, perm1, n, g, keep, nullptr, nullptr, true);
    launch_gather_layout(s, ctx->ph_rec.as<float4>(), ctx->ph_rec.as<float4>() + 1, nullptr, sorted, n,
                         ctx->ph_a.as<float4>(), ctx->ph_b.as<float4>(),
                         ctx->rp.gather_mode == GATHER_LANES ? nullptr : ctx->ph_c.as<float4>(),
                         g.sub_bits > 0 ? ctx->boxes.as<float4>() : nullptr, 2);
    TRY_HIP(ctx, hipGetLastError());
    TRY_HIP(ctx, hipEventRecord(ctx->ev[3], s));
    // gather, once the hit points are sorted on the second stream
    TRY_HIP(ctx, hipStreamWaitEvent(s, ctx->ev[5], 0));
    TRY_HIP(ctx, hipEventRecord(ctx->ev[12], s));
    GatherArgs ga;
    ga.mode = ctx->rp.gather_mode;
    ga.ph_a = ctx->ph_a.as<float4>();
    ga.ph_b = ctx->ph_b.as<float4>();
    ga.ph_c = ctx->ph_c.as<float4>();
    ga.off = ctx->off.as<uint32_t>();
    ga.hp_pos = ctx->hp_pos_s.as<float4>();
    ga.hp_w = ctx->hp_w_s.as<float4>();
    ga.hp_count_src = dc;
    ga.n_hit = npix;
    ga.g = g;
    ga.radius = radius;
    ga.img_sum = ctx->img_sum.as<float>();
    ga.img_comp = ctx->img_comp.as<float>();
    ga.out_rgb = nullptr;
    ga.counters = dc;
    // contribution records need the photons the gather accepts: the lanes
    // kernel marks them as it goes, the others through a separate pass
    ga.boxes = g.sub_bits > 0 ? ctx->boxes.as<float4>() : nullptr;
    const bool mark_in_gather = ctx->prop.enabled && n > 0 && ga.mode == GATHER_LANES;
    ga.used = mark_in_gather ? ctx->prop_used.as<uint32_t>() : nullptr;
    if (mark_in_gather) launch_zero_u32(s, ctx->prop_used.p, n);
    launch_gather(s, ga);
    TRY_HIP(ctx, hipGetLastError());
    TRY_HIP(ctx, hipEventRecord(ctx->ev[6], s));
    if (ctx->prop.enabled && n > 0) {
        // contribution records: deposits that this pass's gather accepted
        if (mark_in_gather) {
            launch_scatter_used(s, ctx->prop_used.as<uint32_t>(), sorted, n, ctx->prop_flags.as<uint32_t>());
        } else {
            launch_zero_u32(s, ctx->prop_flags.p, n);
            launch_mark_used(s, ga, sorted, ctx->prop_flags.as<uint32_t>());
        }
        ContribArgs ca;
        ca.ph_pos = ctx->ph_rec.as<float4>();
        ca.ph_flux = ctx->ph_rec.as<float4>() + 1;
        ca.rec_stride = 2;
        ca.perm = perm1;
        ca.n = n;
        ca.used = ctx->prop_flags.as<uint32_t>();
        ca.seg_c = ctx->prop_segc.as<float>();
        ca.scan = ctx->prop_scan.as<uint64_t>();
        ca.prop = pa.prop;
        ca.seed = ctx->rp.seed;
        ca.pass = (uint32_t)pass;
        ca.scene_id = ctx->rp.scene_id;
        ca.ring = ctx->prop_ring.as<float4>();
        ca.ring_cap = ctx->prop.ring_cap;
        ca.head = ctx->prop.ring_head;
        ca.counters = dc;
        launch_contrib(s, ca, ctx->prop_bsums.as<uint64_t>());
        TRY_HIP(ctx, hipGetLastError());
    }
    TRY_HIP(ctx, hipEventRecord(ctx->ev[7], s));
    if (ctx->sss_present) {
        // eye paths that entered a translucent surface: the dipole sum
        DipoleArgs da;
        da.pts = ctx->sss_pts.as<float4>();
        da.E = ctx->sss_E.as<float>();
        da.n_pts = ctx->sss_n;
        da.area = (float)ctx->sss_area;
        da.dip = ctx->dips.as<DipoleDev>();
        da.hit_pos = ctx->sss_pos.as<float4>();
        da.hit_w = ctx->sss_w.as<float4>();
        da.count_src = dc;
        da.n_hit = npix;
        da.img_sum = ctx->img_sum.as<float>();
        da.img_comp = ctx->img_comp.as<float>();
        da.out_rgb = nullptr;
        launch_dipole(s, da);
        TRY_HIP(ctx, hipGetLastError());
    }
    TRY_HIP(ctx, hipEventRecord(ctx->ev[10], s));
    TRY_HIP(ctx, hipMemcpyAsync(&hc, dc, sizeof(Counters), hipMemcpyDeviceToHost, s));
    TRY_HIP(ctx, hipStreamSynchronize(s));
    if (hc.stack_overflow) return ctx->fail(PPPM_ERR_HIP, "traversal stack overflow");
    double train_ms = 0.0, train_loss = 0.0;
    bool retrained = false;
    if (ctx->prop.enabled) {
        const int64_t recs = (int64_t)hc.proposal_records;
        ctx->prop.ring_head = (ctx->prop.ring_head + recs) % ctx->prop.ring_cap;
        ctx->prop.ring_valid = std::min(ctx->prop.ring_valid + recs, ctx->prop.ring_cap);
        if (pass % ctx->prop.pp.train_interval == 0 && ctx->prop.ring_valid > 0) {
            const auto t0 = std::chrono::steady_clock::now();
            TRY_HIP(ctx, proposal_retrain(ctx, &train_loss));
            train_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
            ctx->prop.last_loss = train_loss;
            retrained = true;
        }
    }
    ctx->passes_done = pass;
    ctx->last_valid = true;
    ctx->last_n = n;
    ctx->last_n_irr = n_irr;
    ctx->last_S = (int32_t)S;
    ctx->last_scale = g.scale;
    ctx->last_perm1 = perm1;
    ctx->last_sorted = sorted;
    ctx->last_sub_bits = g.sub_bits;
    if (st) {
        // stream 1: photon (0-1), order (1-2), grid (2-3), gather (12-6);
        // stream 2: eye (11-4), hit point sort (4-5)
        float t[6] = {0, 0, 0, 0, 0, 0};
        (void)hipEventElapsedTime(&t[0], ctx->ev[0], ctx->ev[1]);
        (void)hipEventElapsedTime(&t[1], ctx->ev[1], ctx->ev[2]);
        (void)hipEventElapsedTime(&t[2], ctx->ev[2], ctx->ev[3]);
        (void)hipEventElapsedTime(&t[3], ctx->ev[11], ctx->ev[4]);
        (void)hipEventElapsedTime(&t[4], ctx->ev[4], ctx->ev[5]);
        (void)hipEventElapsedTime(&t[5], ctx->ev[12], ctx->ev[6]);
        float tot = 0.0f;
        (void)hipEventElapsedTime(&tot, ctx->ev[0], ctx->ev[6]);
        st->hash_size = (int32_t)S;
        st->photon_segments = (int64_t)hc.photon_segments;
        st->eye_segments = (int64_t)hc.eye_segments;
        st->hitpoints = (int64_t)hc.hitpoints;
        st->gather_buckets = (int64_t)hc.gather_buckets;
        st->gather_candidates = (int64_t)hc.gather_candidates;
        st->gather_inradius = (int64_t)hc.gather_inradius;
        st->gather_tested = (int64_t)hc.gather_tested;
        st->gather_blocks = (int64_t)hc.gather_blocks;
        st->photon_nodes = (int64_t)hc.photon_nodes;
        st->photon_tris = (int64_t)hc.photon_tris;
        st->eye_nodes = (int64_t)hc.eye_nodes;
        st->eye_tris = (int64_t)hc.eye_tris;
        st->photon_ms = t[0];
        st->order_ms = t[1];
        st->grid_ms = t[2];
        st->eye_ms = t[3];
        st->hitsort_ms = t[4];
        st->gather_ms = t[5];
        st->total_ms = tot;
        float t_irr = 0.0f, t_dip = 0.0f, t_end = 0.0f;
        (void)hipEventElapsedTime(&t_irr, ctx->ev[8], ctx->ev[9]);
        (void)hipEventElapsedTime(&t_dip, ctx->ev[7], ctx->ev[10]);
        (void)hipEventElapsedTime(&t_end, ctx->ev[0], ctx->ev[10]);
        st->order_ms = t[1] - t_irr;
        st->irradiance_ms = t_irr;
        st->dipole_ms = t_dip;
        st->total_ms = t_end;
        st->irradiance_photons = n_irr;
        st->sss_hits = (int64_t)hc.sss_hits;
        st->sss_points = ctx->sss_n;
        float tp = 0.0f;
        (void)hipEventElapsedTime(&tp, ctx->ev[6], ctx->ev[7]);
        st->proposal_ms = ctx->prop.enabled ? (double)tp + train_ms : 0.0;
        st->train_loss = ctx->prop.last_loss;
        st->proposal_records = (int64_t)hc.proposal_records;
        st->proposal_active = prop_active ? 1 : 0;
        st->retrained = retrained ? 1 : 0;
    }
    return PPPM_OK;
}

// Pipelined passes on four streams:
//   A (stream):  photon pass, then the photon order, hash grid, gather layout;
//   B0 / B1 (stream2 / stream4, even / odd passes): eye pass, hit point sort;
//   C (stream3): the direct emission the eye pass saw, then the gather.
// The eye side of a pass does not depend on its photons, so it is enqueued two
// passes ahead, and consecutive eye passes (whose last long specular paths
// keep few waves busy) overlap on their two streams.  What the C side reads
// (pixel slots, sorted hit points, gather layout, bucket offsets, block boxes,
// counters) rotates over three buffer sets; a set is refilled only after the
// gather three passes back has finished.  C adds every pass's direct emission
// and gather in pass order, so the image equals that of the same number of
// pppm_render_pass calls bit for bit.
namespace {
struct PassSet {
    DBuf *ph_a, *ph_b, *ph_c, *off, *boxes, *hp_pos, *hp_w, *hp_pos_s, *hp_w_s, *counters;
};
}  // namespace

int pppm_render_passes(void* ctx_raw, int32_t count, pppm_pass_stats* stats) {
    CHECK_CTX(ctx);
    if (!ctx->rendering) return ctx->fail(PPPM_ERR_STATE, "pppm_render_begin not called");
    if (count < 0 || ctx->passes_done + count > ctx->rp.passes)
        return ctx->fail(PPPM_ERR_STATE, "more passes than the render was begun with");
    if (count == 0) return PPPM_OK;
    if (ctx->prop.enabled || ctx->prop.active || ctx->sss_present) {
        // the proposal and the subsurface estimate work between passes on the
        // host: one pass at a time
        for (int32_t k = 0; k < count; k++) {
            int rc = pppm_render_pass(ctx_raw, stats ? &stats[k] : nullptr);
            if (rc != PPPM_OK) return rc;
        }
        return PPPM_OK;
    }
    (void)hipSetDevice(ctx->device);
    const int64_t npix = (int64_t)ctx->cam.width * ctx->cam.height;
    PassSet set[3] = {
        {&ctx->ph_a, &ctx->ph_b, &ctx->ph_c, &ctx->off, &ctx->boxes, &ctx->hp_pos, &ctx->hp_w, &ctx->hp_pos_s,
         &ctx->hp_w_s, &ctx->counters},
        {&ctx->ph_a_alt, &ctx->ph_b_alt, &ctx->ph_c_alt, &ctx->off_alt, &ctx->boxes_alt, &ctx->hp_pos_alt,
         &ctx->hp_w_alt, &ctx->hp_pos_s_alt, &ctx->hp_w_s_alt, &ctx->counters_alt},
        {&ctx->ph_a_alt2, &ctx->ph_b_alt2, &ctx->ph_c_alt2, &ctx->off_alt2, &ctx->boxes_alt2, &ctx->hp_pos_alt2,
         &ctx->hp_w_alt2, &ctx->hp_pos_s_alt2, &ctx->hp_w_s_alt2, &ctx->counters_alt2}};
    for (int q = 1; q < 3; q++) {
        TRY_HIP(ctx, set[q].ph_a->ensure(ctx->ph_a.n));
        TRY_HIP(ctx, set[q].ph_b->ensure(ctx->ph_b.n));
        TRY_HIP(ctx, set[q].ph_c->ensure(ctx->ph_c.n));
        TRY_HIP(ctx, set[q].off->ensure(ctx->off.n));
        TRY_HIP(ctx, set[q].boxes->ensure(ctx->boxes.n));
        TRY_HIP(ctx, set[q].hp_pos->ensure(ctx->hp_pos.n));
        TRY_HIP(ctx, set[q].hp_w->ensure(ctx->hp_w.n));
        TRY_HIP(ctx, set[q].hp_pos_s->ensure(ctx->hp_pos_s.n));
        TRY_HIP(ctx, set[q].hp_w_s->ensure(ctx->hp_w_s.n));
        TRY_HIP(ctx, set[q].counters->ensure(ctx->counters.n));
        // every later pass of this render must fit the set
        if (set[q].off->n < ctx->off.n || set[q].ph_a->n < ctx->ph_a.n || set[q].boxes->n < ctx->boxes.n)
            return ctx->fail(PPPM_ERR_NOMEM, "pipelined buffer set");
    }
    for (int i = 0; i < 4; i++) TRY_HIP(ctx, ctx->hk_alt[i].ensure(ctx->hk[i].n));
    TRY_HIP(ctx, ctx->hhist_alt.ensure(ctx->hhist.n));
    TRY_HIP(ctx, ctx->pass_counters.ensure(sizeof(Counters) * (size_t)count));
    hipStream_t A = ctx->stream, Cs = ctx->stream3;
    hipStream_t Bq[2] = {ctx->stream2, ctx->stream4};
    DBuf* hkq[2] = {ctx->hk, ctx->hk_alt};
    DBuf* hhq[2] = {&ctx->hhist, &ctx->hhist_alt};
    enum { ZERO, P0, P1, O0, ODONE, E0, E1, HDONE, G0, GEND, GDONE };
    const int NEV = 6;  // event ring: pass k records into pev[k % NEV]
    const int eye_grid = ctx->cu > 0 ? ctx->cu * EYE_BLOCKS_PER_CU : 2048;
    const int photon_grid = ctx->cu > 0 ? ctx->cu * PHOTON_BLOCKS_PER_CU : 2048;
    const int32_t first = ctx->passes_done + 1;  // pass index of k = 0
    std::vector<Counters> hc_pass((size_t)count);
    std::vector<float> times((size_t)count * 6, 0.0f);
    std::vector<int32_t> sizes((size_t)count, 0);
    std::vector<double> rads((size_t)count, 0.0);
    for (int32_t k = 0; k < count; k++) rads[(size_t)k] = ctx->radii[(size_t)(first + k) - 1];
    std::vector<bool> timed((size_t)count, false);
    auto read_times = [&](int32_t k) {
        if (timed[(size_t)k]) return;
        timed[(size_t)k] = true;
        hipEvent_t* E = ctx->pev[k % NEV];
        (void)hipEventSynchronize(E[GDONE]);
        float* t = &times[(size_t)k * 6];
        (void)hipEventElapsedTime(&t[0], E[P0], E[P1]);
        (void)hipEventElapsedTime(&t[1], E[O0], E[ODONE]);
        (void)hipEventElapsedTime(&t[2], E[E0], E[E1]);
        (void)hipEventElapsedTime(&t[3], E[E1], E[HDONE]);
        (void)hipEventElapsedTime(&t[4], E[G0], E[GEND]);
        (void)hipEventElapsedTime(&t[5], E[ZERO], E[GDONE]);
    };
    // eye side of pass k on B[k % 2]: counters zeroed, eye pass into the pixel
    // slots of set k % 3, hit point sort into the set's sorted buffers
    auto enqueue_eye = [&](int32_t k) -> hipError_t {
        const PassSet& S = set[k % 3];
        hipStream_t B = Bq[k & 1];
        hipEvent_t* E = ctx->pev[k % NEV];
        hipError_t e;
        if (k >= 3 && (e = hipStreamWaitEvent(B, ctx->pev[(k - 3) % NEV][GDONE], 0)) != hipSuccess) return e;
        launch_zero_u32(B, S.counters->p, sizeof(Counters) / 4);
        if ((e = hipEventRecord(E[ZERO], B)) != hipSuccess) return e;
        if ((e = hipEventRecord(E[E0], B)) != hipSuccess) return e;
        EyePassArgs ea = eye_args(ctx, (uint32_t)(first + k));
        ea.hp_pos = S.hp_pos->as<float4>();
        ea.hp_w = S.hp_w->as<float4>();
        ea.counters = S.counters->as<Counters>();
        ea.img_sum = nullptr;  // the direct emission stays in the slots for C
        ea.img_comp = nullptr;
        launch_eye_pass(B, ea, eye_grid);
        if ((e = hipEventRecord(E[E1], B)) != hipSuccess) return e;
        GridParams gh;
        for (int d = 0; d < 3; d++) gh.bmin[d] = ctx->grid_bmin[d];
        gh.scale = 1.0 / (2.0 * rads[(size_t)k]);
        gh.size = 1;
        sort_hitpoints(B, hkq[k & 1], *hhq[k & 1], S.hp_pos->as<float4>(), S.hp_w->as<float4>(), npix, gh,
                       rads[(size_t)k], S.hp_pos_s->as<float4>(), S.hp_w_s->as<float4>());
        if ((e = hipGetLastError()) != hipSuccess) return e;
        return hipEventRecord(E[HDONE], B);
    };
    TRY_HIP(ctx, enqueue_eye(0));
    if (count > 1) TRY_HIP(ctx, enqueue_eye(1));
    for (int32_t k = 0; k < count; k++) {
        // the event set of pass k + 2 was last used by pass k - 4
        if (k >= 3) read_times(k - 3);
        if (k + 2 < count) TRY_HIP(ctx, enqueue_eye(k + 2));
        const PassSet& S = set[k % 3];
        hipEvent_t* E = ctx->pev[k % NEV];
        const int32_t pass = first + k;
        const double radius = rads[(size_t)k];
        Counters* dc = S.counters->as<Counters>();
        // photon side: the counters are zeroed on the eye stream
        TRY_HIP(ctx, hipStreamWaitEvent(A, E[ZERO], 0));
        TRY_HIP(ctx, hipEventRecord(E[P0], A));
        PhotonPassArgs pa;
        pa.sc = ctx->scene_dev();
        pa.n_photons = ctx->rp.photons;
        pa.cap = ctx->cap;
        pa.seed = ctx->rp.seed;
        pa.pass = (uint32_t)pass;
        pa.scene_id = ctx->rp.scene_id;
        pa.max_bounces = ctx->rp.max_bounces;
        pa.ray_eps = (float)ctx->rp.ray_eps;
        pa.ph_pos = ctx->ph_rec.as<float4>();
        pa.ph_flux = ctx->ph_rec.as<float4>() + 1;
        pa.rec_stride = 2;
        pa.counters = dc;
        pa.prop.cdf = nullptr;
        pa.prop.lambda = 1.0f;
        pa.prop.side = 0;
        pa.irr_pos = nullptr;
        pa.irr_flux = nullptr;
        pa.irr_cap = 0;
        pa.count_traversal = ctx->count_traversal;
        pa.rec_rank = ctx->rec_rank.as<uint8_t>();
        pa.rec_count = ctx->rec_count.as<uint32_t>();
        launch_photon_pass(A, pa, photon_grid);
        TRY_HIP(ctx, hipGetLastError());
        TRY_HIP(ctx, hipEventRecord(E[P1], A));
        // the stored count sizes the sorts; meanwhile the eye passes and the
        // previous gather run
        Counters& hc = hc_pass[(size_t)k];
        TRY_HIP(ctx, hipMemcpyAsync(&hc, dc, sizeof(Counters), hipMemcpyDeviceToHost, A));
        TRY_HIP(ctx, hipStreamSynchronize(A));
        const int64_t n = (int64_t)hc.photons_stored;
        if (n > ctx->cap) {
            (void)hipDeviceSynchronize();
            return ctx->fail(PPPM_ERR_CAPACITY, "photon buffer holds " + std::to_string(ctx->cap) +
                                                    " records, the pass stored " + std::to_string(n));
        }
        TRY_HIP(ctx, hipEventRecord(E[O0], A));
        hipStream_t keep_stream = ctx->stream;
        const uint32_t Sz = ctx->rp.hash_size > 0 ? (uint32_t)ctx->rp.hash_size : (uint32_t)(n > 0 ? n : 1);
        sizes[(size_t)k] = (int32_t)Sz;
        GridParams g;
        for (int d = 0; d < 3; d++) g.bmin[d] = ctx->grid_bmin[d];
        g.scale = 1.0 / (2.0 * radius);
        g.size = Sz;
        g.sub_bits = render_sub_bits(ctx, Sz, false);
        const int keep = 4;
        uint32_t* perm1 = origin_order(ctx, A, n, g, nullptr);
        ctx->stream = A;
        uint32_t* sorted = build_grid(ctx, nullptr, nullptr, perm1, n, g, keep, nullptr,
                                      S.off->as<uint32_t>(), true);
        ctx->stream = keep_stream;
        launch_gather_layout(A, ctx->ph_rec.as<float4>(), ctx->ph_rec.as<float4>() + 1, nullptr, sorted, n,
                             S.ph_a->as<float4>(), S.ph_b->as<float4>(),
                             ctx->rp.gather_mode == GATHER_LANES ? nullptr : S.ph_c->as<float4>(),
                             g.sub_bits > 0 ? S.boxes->as<float4>() : nullptr, 2);
        TRY_HIP(ctx, hipGetLastError());
        TRY_HIP(ctx, hipEventRecord(E[ODONE], A));
        // C: direct emission and gather of this pass, in pass order
        TRY_HIP(ctx, hipStreamWaitEvent(Cs, E[ODONE], 0));
        TRY_HIP(ctx, hipStreamWaitEvent(Cs, E[HDONE], 0));
        TRY_HIP(ctx, hipEventRecord(E[G0], Cs));
        launch_direct_add(Cs, S.hp_w->as<float4>(), npix, ctx->img_dsum.as<float>(), ctx->img_dcomp.as<float>());
        GatherArgs ga;
        ga.mode = ctx->rp.gather_mode;
        ga.ph_a = S.ph_a->as<float4>();
        ga.ph_b = S.ph_b->as<float4>();
        ga.ph_c = S.ph_c->as<float4>();
        ga.off = S.off->as<uint32_t>();
        ga.hp_pos = S.hp_pos_s->as<float4>();
        ga.hp_w = S.hp_w_s->as<float4>();
        ga.hp_count_src = dc;
        ga.n_hit = npix;
        ga.g = g;
        ga.radius = radius;
        ga.img_sum = ctx->img_sum.as<float>();
        ga.img_comp = ctx->img_comp.as<float>();
        ga.out_rgb = nullptr;
        ga.counters = dc;
        ga.used = nullptr;
        ga.boxes = g.sub_bits > 0 ? S.boxes->as<float4>() : nullptr;
        launch_gather(Cs, ga);
        TRY_HIP(ctx, hipGetLastError());
        TRY_HIP(ctx, hipEventRecord(E[GEND], Cs));
        TRY_HIP(ctx, hipMemcpyAsync(ctx->pass_counters.as<Counters>() + k, dc, sizeof(Counters),
                                    hipMemcpyDeviceToDevice, Cs));
        TRY_HIP(ctx, hipEventRecord(E[GDONE], Cs));
        ctx->passes_done = pass;
        ctx->last_n = n;
        ctx->last_n_irr = 0;
        ctx->last_S = (int32_t)Sz;
        ctx->last_scale = g.scale;
        ctx->last_perm1 = perm1;
        ctx->last_sorted = sorted;
        ctx->last_sub_bits = g.sub_bits;
    }
    for (int32_t k = std::max(0, count - 3); k < count; k++) read_times(k);
    TRY_HIP(ctx, hipStreamSynchronize(Cs));
    TRY_HIP(ctx, hipStreamSynchronize(Bq[0]));
    TRY_HIP(ctx, hipStreamSynchronize(Bq[1]));
    // the last pass's set becomes the current one (what pppm_last_* read)
    const int last = (count - 1) % 3;
    if (last != 0) {
        auto swap_buf = [](DBuf* a, DBuf* b) {
            std::swap(a->p, b->p);
            std::swap(a->n, b->n);
        };
        swap_buf(set[0].ph_a, set[last].ph_a);
        swap_buf(set[0].ph_b, set[last].ph_b);
        swap_buf(set[0].ph_c, set[last].ph_c);
        swap_buf(set[0].off, set[last].off);
        swap_buf(set[0].boxes, set[last].boxes);
        swap_buf(set[0].hp_pos, set[last].hp_pos);
        swap_buf(set[0].hp_w, set[last].hp_w);
        swap_buf(set[0].hp_pos_s, set[last].hp_pos_s);
        swap_buf(set[0].hp_w_s, set[last].hp_w_s);
        swap_buf(set[0].counters, set[last].counters);
    }
    std::vector<Counters> dcs((size_t)count);
    TRY_HIP(ctx, hipMemcpy(dcs.data(), ctx->pass_counters.p, sizeof(Counters) * (size_t)count, hipMemcpyDeviceToHost));
    for (int32_t k = 0; k < count; k++)
        if (dcs[(size_t)k].stack_overflow) return ctx->fail(PPPM_ERR_HIP, "traversal stack overflow");
    ctx->last_valid = true;
    if (stats) {
        for (int32_t k = 0; k < count; k++) {
            pppm_pass_stats& st = stats[k];
            const Counters& c = dcs[(size_t)k];
            const float* t = &times[(size_t)k * 6];
            std::memset(&st, 0, sizeof(st));
            st.pass_index = first + k;
            st.hash_size = sizes[(size_t)k];
            st.radius = rads[(size_t)k];
            st.photons_emitted = ctx->rp.photons;
            st.photons_stored = (int64_t)c.photons_stored;
            st.photon_segments = (int64_t)c.photon_segments;
            st.eye_segments = (int64_t)c.eye_segments;
            st.hitpoints = (int64_t)c.hitpoints;
            st.gather_buckets = (int64_t)c.gather_buckets;
            st.gather_candidates = (int64_t)c.gather_candidates;
            st.gather_inradius = (int64_t)c.gather_inradius;
            st.gather_tested = (int64_t)c.gather_tested;
            st.gather_blocks = (int64_t)c.gather_blocks;
            st.photon_nodes = (int64_t)c.photon_nodes;
            st.photon_tris = (int64_t)c.photon_tris;
            st.eye_nodes = (int64_t)c.eye_nodes;
            st.eye_tris = (int64_t)c.eye_tris;
            st.photon_ms = t[0];
            st.order_ms = t[1];
            st.grid_ms = 0.0;
            st.eye_ms = t[2];
            st.hitsort_ms = t[3];
            st.gather_ms = t[4];
            st.total_ms = t[5];
        }
    }
    return PPPM_OK;
}

int pppm_count_traversal(void* ctx_raw, int32_t enable) {
    CHECK_CTX(ctx);
    ctx->count_traversal = enable ? 1 : 0;
    return PPPM_OK;
}

int pppm_image_read(void* ctx_raw, double* rgb, int64_t n_values) {
    CHECK_CTX(ctx);
    if (!ctx->rendering) return ctx->fail(PPPM_ERR_STATE, "no render in progress");
    const int64_t nv = 3 * (int64_t)ctx->cam.width * ctx->cam.height;
    if (!rgb || n_values < nv) return ctx->fail(PPPM_ERR_INVALID, "image buffer too small");
    (void)hipSetDevice(ctx->device);
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
    if (gather_mode < PPPM_GATHER_AUTO || gather_mode > PPPM_GATHER_WAVE)
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
    // (the lanes kernel's reads past the sorted positions p4s land in f4s)
    size_t bytes = sizeof(float) * 9 * (size_t)(np + nh) + sizeof(float4) * 5 * (size_t)np +
                   sizeof(float4) * 2 * (size_t)nh + sizeof(float) * 3 * (size_t)nh;
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
    if (gather_mode == GATHER_LANES)
        g.sub_bits = std::max(0, std::min(3, (32 - grid_bucket_bits(g.size, false)) / 3));
    uint32_t* sorted = build_grid(ctx, p4, nullptr, nullptr, n_ph, g, -1);
    launch_gather_layout(s, p4, f4, nullptr, sorted, n_ph, p4s, f4s, gather_mode == GATHER_LANES ? nullptr : c4s);
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

int pppm_subsurface_configure(void* ctx_raw, int64_t points) {
    CHECK_CTX(ctx);
    if (points < 0 || points > (1LL << 26)) return ctx->fail(PPPM_ERR_INVALID, "points: 0 (default) .. 2^26");
    ctx->sss_points_req = points;
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
        if (n_pts > 0) {
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

int pppm_proposal_enable(void* ctx_raw, const pppm_proposal_params* p) {
    CHECK_CTX(ctx);
    if (!ctx->rendering) return ctx->fail(PPPM_ERR_STATE, "call pppm_render_begin first");
    if (!p || !(p->lambda > 0.0 && p->lambda <= 1.0) || !(p->learning_rate >= 0.0) || p->anchors < 1 ||
        p->anchors > 32 || p->train_interval < 1 || p->train_steps < 1 || p->batch < 256 || p->batch % 256 != 0 ||
        p->batch > 65536 || p->ring_capacity < 1 || p->ring_capacity > (1LL << 28))
        return ctx->fail(PPPM_ERR_INVALID, "bad proposal parameters");
    if (ctx->h_lights.size() != 1) return ctx->fail(PPPM_ERR_INVALID, "the learned proposal needs exactly one light");
    (void)hipSetDevice(ctx->device);
    ctx->prop.pp = *p;
    std::vector<float> w = mlp_init_params(p->seed);
    TRY_HIP(ctx, mlp_upload(ctx, w.data()));
    TRY_HIP(ctx, ensure_train(ctx, p->batch));
    const int64_t cap = ctx->cap > 0 ? ctx->cap : 1;
    TRY_HIP(ctx, ctx->prop_flags.ensure(sizeof(uint32_t) * (size_t)cap));
    TRY_HIP(ctx, ctx->prop_used.ensure(sizeof(uint32_t) * (size_t)cap));
    TRY_HIP(ctx, ctx->prop_segc.ensure(sizeof(float) * (size_t)cap));
    TRY_HIP(ctx, ctx->prop_scan.ensure(sizeof(uint64_t) * (size_t)cap));
    TRY_HIP(ctx, ctx->prop_bsums.ensure(sizeof(uint64_t) * (size_t)(cap / 1024 + cap / (1024 * 512) + 64)));
    TRY_HIP(ctx, ctx->prop_ring.ensure(sizeof(float4) * (size_t)p->ring_capacity));
    TRY_HIP(ctx, hipStreamSynchronize(ctx->stream));
    ctx->prop.ring_cap = p->ring_capacity;
    ctx->prop.ring_head = 0;
    ctx->prop.ring_valid = 0;
    ctx->prop.draws = 0;
    ctx->prop.last_loss = 0.0;
    ctx->prop.enabled = true;
    ctx->prop.active = false;
    return PPPM_OK;
}

int pppm_proposal_set_table(void* ctx_raw, const float* prob, int32_t anchors, double lambda) {
    CHECK_CTX(ctx);
    if (!ctx->rendering) return ctx->fail(PPPM_ERR_STATE, "call pppm_render_begin first");
    if (!prob || anchors < 1 || anchors > 32 || !(lambda > 0.0 && lambda <= 1.0))
        return ctx->fail(PPPM_ERR_INVALID, "bad proposal table");
    if (ctx->h_lights.size() != 1) return ctx->fail(PPPM_ERR_INVALID, "the learned proposal needs exactly one light");
    (void)hipSetDevice(ctx->device);
    const size_t A = (size_t)anchors * anchors;
    TRY_HIP(ctx, ctx->prop_prob.ensure(sizeof(float) * PROP_CELLS * A));
    TRY_HIP(ctx, ctx->prop_cdf.ensure(sizeof(float) * PROP_CELLS * A));
    TRY_HIP(ctx, hipMemcpyAsync(ctx->prop_prob.p, prob, sizeof(float) * PROP_CELLS * A, hipMemcpyHostToDevice,
                                ctx->stream));
    launch_prop_table(ctx->stream, ctx->prop_prob.as<float>(), (int32_t)A, ctx->prop_cdf.as<float>());
    TRY_HIP(ctx, hipGetLastError());
    TRY_HIP(ctx, hipStreamSynchronize(ctx->stream));
    ctx->prop.side = anchors;
    ctx->prop.lambda = (float)lambda;
    ctx->prop.active = true;
    return PPPM_OK;
}

int pppm_proposal_get_table(void* ctx_raw, float* cdf, int64_t n_values, int32_t* anchors) {
    CHECK_CTX(ctx);
    if (!ctx->prop.active) return ctx->fail(PPPM_ERR_STATE, "no proposal table");
    const int64_t nv = (int64_t)ctx->prop.side * ctx->prop.side * PROP_CELLS;
    if (anchors) *anchors = ctx->prop.side;
    if (!cdf) return PPPM_OK;
    if (n_values < nv) return ctx->fail(PPPM_ERR_INVALID, "table buffer too small");
    (void)hipSetDevice(ctx->device);
    TRY_HIP(ctx, hipMemcpy(cdf, ctx->prop_cdf.p, sizeof(float) * (size_t)nv, hipMemcpyDeviceToHost));
    return PPPM_OK;
}

int pppm_proposal_records(void* ctx_raw, float* rec, int64_t cap, int64_t* n) {
    CHECK_CTX(ctx);
    if (!ctx->prop.enabled) return ctx->fail(PPPM_ERR_STATE, "the proposal is not enabled");
    const int64_t nv = ctx->prop.ring_valid;
    if (n) *n = nv;
    if (!rec || cap <= 0) return PPPM_OK;
    (void)hipSetDevice(ctx->device);
    std::vector<float4> all((size_t)nv);
    if (nv > 0)
        TRY_HIP(ctx, hipMemcpy(all.data(), ctx->prop_ring.p, sizeof(float4) * (size_t)nv, hipMemcpyDeviceToHost));
    // oldest first: a full ring starts at the head
    const int64_t start = nv == ctx->prop.ring_cap ? ctx->prop.ring_head : 0;
    const int64_t m = std::min(nv, cap);
    for (int64_t i = 0; i < m; i++) {
        const float4 r = all[(size_t)((start + i) % (nv > 0 ? nv : 1))];
        rec[4 * i] = r.x;
        rec[4 * i + 1] = r.y;
        uint32_t cell;
        std::memcpy(&cell, &r.z, 4);
        rec[4 * i + 2] = (float)cell;
        rec[4 * i + 3] = r.w;
    }
    return PPPM_OK;
}

int pppm_mlp_set_params(void* ctx_raw, const float* params, int64_t n) {
    CHECK_CTX(ctx);
    if (!params || n != MLP_NPARAM) return ctx->fail(PPPM_ERR_INVALID, "expected PPPM_MLP_PARAMS values");
    (void)hipSetDevice(ctx->device);
    TRY_HIP(ctx, mlp_upload(ctx, params));
    TRY_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return PPPM_OK;
}

int pppm_mlp_get_params(void* ctx_raw, float* params, int64_t n) {
    CHECK_CTX(ctx);
    if (!params || n < MLP_NPARAM) return ctx->fail(PPPM_ERR_INVALID, "expected PPPM_MLP_PARAMS values");
    (void)hipSetDevice(ctx->device);
    TRY_HIP(ctx, ensure_mlp(ctx));
    TRY_HIP(ctx, hipMemcpy(params, ctx->mlp_params.p, sizeof(float) * MLP_NPARAM, hipMemcpyDeviceToHost));
    return PPPM_OK;
}

int pppm_mlp_forward(void* ctx_raw, const float* uv, int64_t n, float* prob) {
    CHECK_CTX(ctx);
    if (n < 0 || (n > 0 && (!uv || !prob)) || n > 65536) return ctx->fail(PPPM_ERR_INVALID, "bad arguments");
    if (n == 0) return PPPM_OK;
    (void)hipSetDevice(ctx->device);
    TRY_HIP(ctx, ensure_mlp(ctx));
    TRY_HIP(ctx, ctx->mlp_stage.ensure(sizeof(float) * 2 * (size_t)n));
    TRY_HIP(ctx, ctx->scratch.ensure(sizeof(float) * PROP_CELLS * (size_t)n));
    TRY_HIP(ctx, hipMemcpyAsync(ctx->mlp_stage.p, uv, sizeof(float) * 2 * (size_t)n, hipMemcpyHostToDevice,
                                ctx->stream));
    MlpArgs m{};
    m.params = ctx->mlp_params.as<float>();
    m.wb = ctx->mlp_wb.as<uint16_t>();
    m.uv = ctx->mlp_stage.as<float2>();
    m.n = n;
    m.prob = ctx->scratch.as<float>();
    launch_mlp(ctx->stream, m, false);
    TRY_HIP(ctx, hipGetLastError());
    TRY_HIP(ctx, hipMemcpyAsync(prob, ctx->scratch.p, sizeof(float) * PROP_CELLS * (size_t)n, hipMemcpyDeviceToHost,
                                ctx->stream));
    TRY_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return PPPM_OK;
}

int pppm_mlp_train_step(void* ctx_raw, const float* rec, int64_t n, double learning_rate, int32_t step,
                        float* grad, double* loss) {
    CHECK_CTX(ctx);
    if (!rec || n < 256 || n % 256 != 0 || n > 65536 || !(learning_rate >= 0.0))
        return ctx->fail(PPPM_ERR_INVALID, "records: a multiple of 256, at most 65536");
    (void)hipSetDevice(ctx->device);
    TRY_HIP(ctx, ensure_mlp(ctx));
    TRY_HIP(ctx, ensure_train(ctx, n));
    std::vector<float4> r((size_t)n);
    for (int64_t i = 0; i < n; i++) {
        const float c = rec[4 * i + 2];
        if (!(c >= 0.0f && c < (float)PROP_CELLS)) return ctx->fail(PPPM_ERR_INVALID, "cell out of range");
        const uint32_t cell = (uint32_t)c;
        float cell_bits;
        std::memcpy(&cell_bits, &cell, 4);
        r[(size_t)i] = make_float4(rec[4 * i], rec[4 * i + 1], cell_bits, rec[4 * i + 3]);
    }
    TRY_HIP(ctx, ctx->mlp_stage.ensure(sizeof(float4) * (size_t)n));
    TRY_HIP(ctx, hipMemcpyAsync(ctx->mlp_stage.p, r.data(), sizeof(float4) * (size_t)n, hipMemcpyHostToDevice,
                                ctx->stream));
    launch_zero_u32(ctx->stream, ctx->mlp_loss.p, 1);
    ctx->prop.adam_t = step;
    mlp_step(ctx, ctx->mlp_stage.as<float4>(), n, n, true, learning_rate);
    TRY_HIP(ctx, hipGetLastError());
    float l = 0.0f;
    TRY_HIP(ctx, hipMemcpyAsync(&l, ctx->mlp_loss.p, sizeof(float), hipMemcpyDeviceToHost, ctx->stream));
    if (grad)
        TRY_HIP(ctx, hipMemcpyAsync(grad, ctx->mlp_grad.p, sizeof(float) * MLP_NPARAM, hipMemcpyDeviceToHost,
                                    ctx->stream));
    TRY_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if (loss) *loss = l;
    return PPPM_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Batched scenes: many renders in one hipGraph-captured pass loop.  No
// reference counterpart (the reference renders one scene per process,
// sources/main/appmain.cc); each scene's pass is the same computation as
// pppm_render_pass with a fixed hash size, enqueued without host
// synchronisation so that one capture holds a pass of every scene.
// ---------------------------------------------------------------------------
namespace {

// Device pass state of a render whose passes are enqueued by enqueue_pass
hipError_t prepare_enqueued(Ctx* ctx) {
    hipError_t e;
    if ((e = ctx->pdev.ensure(sizeof(PassDev))) != hipSuccess) return e;
    if ((e = ctx->radii_dev.ensure(sizeof(double) * ctx->radii.size())) != hipSuccess) return e;
    if ((e = hipMemcpy(ctx->radii_dev.p, ctx->radii.data(), sizeof(double) * ctx->radii.size(),
                       hipMemcpyHostToDevice)) != hipSuccess)
        return e;
    PassDev pd;
    std::memset(&pd, 0, sizeof(pd));
    pd.pass = (uint32_t)ctx->passes_done;
    return hipMemcpy(ctx->pdev.p, &pd, sizeof(pd), hipMemcpyHostToDevice);
}

// One pass of a plain render on ctx->stream with no host synchronisation:
// counts stay on the device, every sort runs over the buffer capacity with
// sentinel keys behind the records, and the pass index and radius come from
// the device pass state (advanced first).  The image equals pppm_render_pass's
// bit for bit for the same hash size.
void enqueue_pass(Ctx* ctx, int photon_grid) {
    hipStream_t s = ctx->stream;
    PassDev* pd = ctx->pdev.as<PassDev>();
    Counters* dc = ctx->counters.as<Counters>();
    const int64_t cap = ctx->cap;
    const int64_t npix = (int64_t)ctx->cam.width * ctx->cam.height;
    launch_pass_advance(s, pd, ctx->radii_dev.as<double>());
    launch_zero_u32(s, dc, sizeof(Counters) / 4);
    PhotonPassArgs pa;
    pa.sc = ctx->scene_dev();
    pa.n_photons = ctx->rp.photons;
    pa.cap = cap;
    pa.seed = ctx->rp.seed;
    pa.pass = 0;
    pa.scene_id = ctx->rp.scene_id;
    pa.max_bounces = ctx->rp.max_bounces;
    pa.ray_eps = (float)ctx->rp.ray_eps;
    pa.ph_pos = ctx->ph_rec.as<float4>();
    pa.ph_flux = ctx->ph_rec.as<float4>() + 1;
    pa.rec_stride = 2;
    pa.counters = dc;
    pa.pd = pd;
    pa.prop.cdf = nullptr;
    pa.prop.lambda = 1.0f;
    pa.prop.side = 0;
    pa.irr_pos = nullptr;
    pa.irr_flux = nullptr;
    pa.irr_cap = 0;
    pa.rec_rank = ctx->rec_rank.as<uint8_t>();
    pa.rec_count = ctx->rec_count.as<uint32_t>();
    launch_photon_pass(s, pa, photon_grid);
    launch_capacity_check(s, pd, dc, cap);
    GridParams g;
    for (int d = 0; d < 3; d++) g.bmin[d] = ctx->grid_bmin[d];
    g.scale = 0.0;
    g.size = (uint32_t)ctx->rp.hash_size;
    g.sub_bits = render_sub_bits(ctx, g.size, true);
    g.pd = pd;
    const int keep = 4;
    const uint32_t* perm1 = origin_order(ctx, s, cap, g, &dc->photons_stored);
    const uint32_t* sorted = build_grid(ctx, nullptr, nullptr, perm1, cap, g, keep,
                                        &dc->photons_stored, nullptr, true);
    launch_gather_layout(s, ctx->ph_rec.as<float4>(), ctx->ph_rec.as<float4>() + 1, nullptr, sorted, cap,
                         ctx->ph_a.as<float4>(), ctx->ph_b.as<float4>(),
                         ctx->rp.gather_mode == GATHER_LANES ? nullptr : ctx->ph_c.as<float4>(),
                         g.sub_bits > 0 ? ctx->boxes.as<float4>() : nullptr, 2);
    EyePassArgs ea = eye_args(ctx, 0);
    ea.pd = pd;
    launch_eye_pass(s, ea, ctx->cu > 0 ? ctx->cu * EYE_BLOCKS_PER_CU : 2048);
    sort_hitpoints(ctx, ctx->stream, ctx->hp_pos.as<float4>(), ctx->hp_w.as<float4>(), npix, g, 0.0);
    GatherArgs ga;
    std::memset(&ga, 0, sizeof(ga));
    ga.mode = ctx->rp.gather_mode;
    ga.ph_a = ctx->ph_a.as<float4>();
    ga.ph_b = ctx->ph_b.as<float4>();
    ga.ph_c = ctx->ph_c.as<float4>();
    ga.off = ctx->off.as<uint32_t>();
    ga.hp_pos = ctx->hp_pos_s.as<float4>();
    ga.hp_w = ctx->hp_w_s.as<float4>();
    ga.hp_count_src = dc;
    ga.n_hit = npix;
    ga.g = g;
    ga.radius = 0.0;
    ga.img_sum = ctx->img_sum.as<float>();
    ga.img_comp = ctx->img_comp.as<float>();
    ga.counters = dc;
    ga.boxes = g.sub_bits > 0 ? ctx->boxes.as<float4>() : nullptr;
    launch_gather(s, ga);
}

struct BatchScene {
    Ctx* ctx = nullptr;
    int photon_grid = 0;     // photon kernel grid (autotuned)
    double cost_ms = 0.0;    // measured pass time of the chosen configuration
    int branch = 0;
};

struct Batch {
    int device = 0;
    int branches = 1;
    std::vector<BatchScene> scenes;
    hipStream_t main = nullptr;
    std::vector<hipStream_t> br;
    std::vector<hipEvent_t> ev;   // [0] fork, [1..] joins, then timing
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    int32_t passes_done = 0;
    std::string err;

    int fail(int code, const std::string& m) {
        err = m;
        return code;
    }
    void drop_graph() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        exec = nullptr;
        graph = nullptr;
    }
};

thread_local std::string g_batch_error;

}  // namespace

extern "C" {

int pppm_batch_create(void** out, int32_t device, int32_t branches) {
    if (!out || branches < 1 || branches > 16) return PPPM_ERR_INVALID;
    *out = nullptr;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || device < 0 || device >= count) {
        g_batch_error = "no such HIP device";
        return PPPM_ERR_HIP;
    }
    (void)hipSetDevice(device);
    Batch* b = new Batch();
    b->device = device;
    b->branches = branches;
    hipError_t e = hipStreamCreate(&b->main);
    b->br.assign((size_t)branches, nullptr);
    for (int i = 0; i < branches && e == hipSuccess; i++) e = hipStreamCreate(&b->br[(size_t)i]);
    b->ev.assign((size_t)branches + 3, nullptr);
    for (auto& x : b->ev)
        if (e == hipSuccess) e = hipEventCreate(&x);
    if (e != hipSuccess) {
        g_batch_error = std::string("stream set-up: ") + hipGetErrorString(e);
        for (auto& x : b->ev)
            if (x) (void)hipEventDestroy(x);
        for (auto& x : b->br)
            if (x) (void)hipStreamDestroy(x);
        if (b->main) (void)hipStreamDestroy(b->main);
        delete b;
        return PPPM_ERR_HIP;
    }
    *out = b;
    return PPPM_OK;
}

const char* pppm_batch_last_error(const void* batch) {
    const Batch* b = reinterpret_cast<const Batch*>(batch);
    return b ? b->err.c_str() : g_batch_error.c_str();
}

int pppm_batch_destroy(void* batch) {
    Batch* b = reinterpret_cast<Batch*>(batch);
    if (!b) return PPPM_ERR_INVALID;
    (void)hipSetDevice(b->device);
    (void)hipStreamSynchronize(b->main);
    b->drop_graph();
    for (auto& sc : b->scenes) pppm_destroy(sc.ctx);
    for (auto& x : b->ev) (void)hipEventDestroy(x);
    for (auto& x : b->br) (void)hipStreamDestroy(x);
    (void)hipStreamDestroy(b->main);
    delete b;
    return PPPM_OK;
}

int pppm_batch_add(void* batch, const float* tri_xyz, const int32_t* bsdf_id, int64_t n, const pppm_bsdf* table,
                   int32_t nb, const pppm_light* lights, int32_t nl, const pppm_camera* cam,
                   const pppm_render_params* params, int32_t* index) {
    Batch* b = reinterpret_cast<Batch*>(batch);
    if (!b) return PPPM_ERR_INVALID;
    if (!params || params->hash_size <= 0) return b->fail(PPPM_ERR_INVALID, "batched scenes need a fixed hash_size > 0");
    for (int32_t i = 0; i < nb; i++)
        if (table && (table[i].type & PPPM_BSDF_BSSRDF))
            return b->fail(PPPM_ERR_INVALID, "batched scenes are plain PPM (no BSSRDF)");
    void* h = nullptr;
    int rc = pppm_create(&h, b->device);
    if (rc != PPPM_OK) return b->fail(rc, pppm_last_error(nullptr));
    Ctx* c = reinterpret_cast<Ctx*>(h);
    pppm_bvh_info info;
    if ((rc = pppm_scene_upload(h, tri_xyz, bsdf_id, n, table, nb, lights, nl)) != PPPM_OK ||
        (rc = pppm_bvh_build(h, &info)) != PPPM_OK || (rc = pppm_camera_set(h, cam)) != PPPM_OK ||
        (rc = pppm_render_begin(h, params)) != PPPM_OK) {
        b->fail(rc, c->err);
        pppm_destroy(h);
        return rc;
    }
    hipError_t e = prepare_enqueued(c);
    if (e != hipSuccess) {
        b->fail(PPPM_ERR_HIP, hipGetErrorString(e));
        pppm_destroy(h);
        return PPPM_ERR_HIP;
    }
    BatchScene sc;
    sc.ctx = c;
    sc.photon_grid = c->cu > 0 ? c->cu * PHOTON_BLOCKS_PER_CU : 2048;
    b->scenes.push_back(sc);
    b->drop_graph();
    if (index) *index = (int32_t)b->scenes.size() - 1;
    return PPPM_OK;
}

// Per scene: every gather kernel x photon-kernel grid on one pass each,
// keeping the fastest pass; then the render restarts and the scenes are
// spread over the branches longest first (LPT).
int pppm_batch_tune(void* batch, int32_t tune_gather, int32_t tune_grid) {
    Batch* b = reinterpret_cast<Batch*>(batch);
    if (!b) return PPPM_ERR_INVALID;
    (void)hipSetDevice(b->device);
    for (auto& sc : b->scenes) {
        Ctx* c = sc.ctx;
        const pppm_render_params base = c->rp;
        std::vector<int> modes;
        if (tune_gather) modes = {GATHER_LANES, GATHER_TILED, GATHER_WAVE};
        else modes = {base.gather_mode};
        std::vector<int> grids;
        const int cu = c->cu > 0 ? c->cu : 256;
        if (tune_grid) grids = {cu, cu * 2, cu * 4, cu * 8};
        else grids = {sc.photon_grid};
        double best = 1e30;
        int best_mode = base.gather_mode, best_grid = sc.photon_grid;
        for (int m : modes) {
            pppm_render_params p = base;
            p.gather_mode = m;
            p.passes = (int32_t)grids.size() + 1;
            int rc = pppm_render_begin(c, &p);
            if (rc != PPPM_OK) return b->fail(rc, c->err);
            if (prepare_enqueued(c) != hipSuccess) return b->fail(PPPM_ERR_HIP, "tuning set-up");
            // first pass warms the caches; then one pass per grid
            enqueue_pass(c, grids[0]);
            for (int gsz : grids) {
                (void)hipEventRecord(c->ev[0], c->stream);
                enqueue_pass(c, gsz);
                (void)hipEventRecord(c->ev[1], c->stream);
                if (hipEventSynchronize(c->ev[1]) != hipSuccess) return b->fail(PPPM_ERR_HIP, "tuning pass");
                float ms = 0.0f;
                (void)hipEventElapsedTime(&ms, c->ev[0], c->ev[1]);
                if (ms < best) {
                    best = ms;
                    best_mode = m;
                    best_grid = gsz;
                }
            }
        }
        pppm_render_params p = base;
        p.gather_mode = best_mode;
        int rc = pppm_render_begin(c, &p);
        if (rc != PPPM_OK) return b->fail(rc, c->err);
        if (prepare_enqueued(c) != hipSuccess) return b->fail(PPPM_ERR_HIP, "tuning reset");
        sc.photon_grid = best_grid;
        sc.cost_ms = best;
    }
    b->drop_graph();
    b->passes_done = 0;
    return PPPM_OK;
}

int pppm_batch_run(void* batch, int32_t passes, double* ms) {
    Batch* b = reinterpret_cast<Batch*>(batch);
    if (!b) return PPPM_ERR_INVALID;
    if (passes < 0) return b->fail(PPPM_ERR_INVALID, "passes < 0");
    if (b->scenes.empty()) return b->fail(PPPM_ERR_STATE, "no scenes");
    for (auto& sc : b->scenes)
        if (sc.ctx->passes_done + passes > sc.ctx->rp.passes)
            return b->fail(PPPM_ERR_STATE, "more passes than the scenes were begun with");
    (void)hipSetDevice(b->device);
    if (!b->exec) {
        // branches longest first: each scene to the least loaded branch
        std::vector<size_t> order(b->scenes.size());
        for (size_t i = 0; i < order.size(); i++) order[i] = i;
        std::stable_sort(order.begin(), order.end(),
                         [&](size_t x, size_t y) { return b->scenes[x].cost_ms > b->scenes[y].cost_ms; });
        std::vector<double> load((size_t)b->branches, 0.0);
        for (size_t i : order) {
            size_t best = 0;
            for (size_t k = 1; k < load.size(); k++)
                if (load[k] < load[best]) best = k;
            b->scenes[i].branch = (int)best;
            load[best] += b->scenes[i].cost_ms > 0.0 ? b->scenes[i].cost_ms : 1.0;
        }
        hipError_t e = hipStreamBeginCapture(b->main, hipStreamCaptureModeThreadLocal);
        if (e != hipSuccess) return b->fail(PPPM_ERR_HIP, std::string("capture: ") + hipGetErrorString(e));
        (void)hipEventRecord(b->ev[0], b->main);
        for (int k = 0; k < b->branches; k++) (void)hipStreamWaitEvent(b->br[(size_t)k], b->ev[0], 0);
        for (size_t i : order) {
            BatchScene& sc = b->scenes[i];
            hipStream_t own = sc.ctx->stream;
            sc.ctx->stream = b->br[(size_t)sc.branch];
            enqueue_pass(sc.ctx, sc.photon_grid);
            sc.ctx->stream = own;
        }
        for (int k = 0; k < b->branches; k++) {
            (void)hipEventRecord(b->ev[(size_t)k + 1], b->br[(size_t)k]);
            (void)hipStreamWaitEvent(b->main, b->ev[(size_t)k + 1], 0);
        }
        e = hipStreamEndCapture(b->main, &b->graph);
        if (e == hipSuccess) e = hipGetLastError();
        if (e == hipSuccess) e = hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0);
        if (e != hipSuccess) {
            b->drop_graph();
            return b->fail(PPPM_ERR_HIP, std::string("graph: ") + hipGetErrorString(e));
        }
    }
    hipEvent_t t0 = b->ev[(size_t)b->branches + 1], t1 = b->ev[(size_t)b->branches + 2];
    (void)hipEventRecord(t0, b->main);
    for (int p = 0; p < passes; p++) {
        hipError_t e = hipGraphLaunch(b->exec, b->main);
        if (e != hipSuccess) return b->fail(PPPM_ERR_HIP, std::string("graph launch: ") + hipGetErrorString(e));
    }
    (void)hipEventRecord(t1, b->main);
    hipError_t e = hipEventSynchronize(t1);
    if (e != hipSuccess) return b->fail(PPPM_ERR_HIP, std::string("batch: ") + hipGetErrorString(e));
    float t = 0.0f;
    (void)hipEventElapsedTime(&t, t0, t1);
    if (ms) *ms = t;
    for (auto& sc : b->scenes) {
        sc.ctx->passes_done += passes;
        PassDev pd;
        if (hipMemcpy(&pd, sc.ctx->pdev.p, sizeof(pd), hipMemcpyDeviceToHost) != hipSuccess)
            return b->fail(PPPM_ERR_HIP, "pass state read-back");
        if (pd.overflow)
            return b->fail(PPPM_ERR_CAPACITY, "a batched scene stored more photons than its capacity");
    }
    b->passes_done += passes;
    return PPPM_OK;
}

int pppm_batch_image(void* batch, int32_t index, double* rgb, int64_t n_values) {
    Batch* b = reinterpret_cast<Batch*>(batch);
    if (!b) return PPPM_ERR_INVALID;
    if (index < 0 || index >= (int32_t)b->scenes.size()) return b->fail(PPPM_ERR_INVALID, "scene index");
    int rc = pppm_image_read(b->scenes[(size_t)index].ctx, rgb, n_values);
    if (rc != PPPM_OK) b->fail(rc, b->scenes[(size_t)index].ctx->err);
    return rc;
}

int pppm_batch_scene_info(void* batch, int32_t index, int32_t* gather_mode, int32_t* photon_grid, int32_t* branch,
                          double* cost_ms) {
    Batch* b = reinterpret_cast<Batch*>(batch);
    if (!b) return PPPM_ERR_INVALID;
    if (index < 0 || index >= (int32_t)b->scenes.size()) return b->fail(PPPM_ERR_INVALID, "scene index");
    const BatchScene& sc = b->scenes[(size_t)index];
    if (gather_mode) *gather_mode = sc.ctx->rp.gather_mode;
    if (photon_grid) *photon_grid = sc.photon_grid;
    if (branch) *branch = sc.branch;
    if (cost_ms) *cost_ms = sc.cost_ms;
    return PPPM_OK;
}

}  // extern "C"
