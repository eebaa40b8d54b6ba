// ctx.h -- host-side state of a renderer context (the object behind the
// void* ctx of include/pppm.h) and the helpers its translation units share:
// capi.hip (context, scene, BVH, grid, inspection and measurement hooks),
// pass.hip (the pass loop), sss_host.hip (subsurface scattering),
// proposal_host.hip (learned emission proposal), batch.hip (batched scenes).
#pragma once

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "internal.h"
#include "traverse.h"

namespace pppmi {


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

// the octree cut of the dipole sum on the device (dipole_cut.hip)
struct CutBufs {
    DBuf a, b, lo, hi, perm, levels;
    std::vector<int64_t> level_off;  // nodes of depth d: levels[level_off[d] .. level_off[d + 1])
    int32_t n_nodes = 0;
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
    // pinned host copies of the pipelined passes' counters, so that their
    // copy does not hold the host while the photon kernel runs
    Counters* pinned_counters = nullptr;
    int64_t pinned_count = 0;
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
    double sss_cut = 0.0;              // 0: the exact sum; > 0: the octree cut's maximum error
    CutBufs cut;                       // the cut over the surface points
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

    // environment map of the scene (pppm_envmap_set): texels, importance cells,
    // and the per-render emission constants (render_begin)
    bool has_env = false;
    int32_t env_w = 0, env_h = 0, env_last = 0;
    std::vector<float> h_env;          // [h][w][3]
    std::vector<double> h_env_pdf, h_env_cdf;
    DBuf env_tex, env_pdf, env_cdf;
    double env_k = 0.0, pick_total = 0.0, pick_area = 0.0;
    // render points of the Hachisuka method: radius^2, photon count, flux
    DBuf ppm_r2, ppm_n, ppm_flux, ppm_bounds, ppm_image;
    // device address of the no-allocation scratch array on this device
    // (pppm_hashgrid_build_noalloc), or null with the lookup's error
    void* noalloc_scratch = nullptr;
    hipError_t noalloc_err = hipSuccess;

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
        s.env = has_env ? env_tex.as<float4>() : nullptr;
        s.env_pdf = env_pdf.as<double>();
        s.env_cdf = env_cdf.as<double>();
        s.env_w = env_w;
        s.env_h = env_h;
        s.env_last = env_last;
        s.env_c = mk((float)(0.5 * (bmin[0] + bmax[0])), (float)(0.5 * (bmin[1] + bmax[1])),
                     (float)(0.5 * (bmin[2] + bmax[2])));
        s.env_r = (float)(0.5 * std::sqrt((bmax[0] - bmin[0]) * (bmax[0] - bmin[0]) +
                                          (bmax[1] - bmin[1]) * (bmax[1] - bmin[1]) +
                                          (bmax[2] - bmin[2]) * (bmax[2] - bmin[2])));
        s.env_k = env_k;
        s.pick_total = pick_total;
        s.pick_area = pick_area;
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

// ---- shared helpers ----
int bits_for(uint64_t max_value);
// host copies of the device octahedral encode / decode (pppm_device.h); the
// encode returns the float whose bits hold the code
float oct_encode_host(v3f n);
void oct_decode_host(uint32_t e, float out[3]);
void cross3(const double a[3], const double b[3], double o[3]);
// Bits of a bucket id in the grid sort key (with sentinel: one more value for
// the sentinel of unused slots).
int grid_bucket_bits(uint32_t size, bool sentinel);
// Cell-local Morton bits per axis below the bucket in the render path's grid
// sort (lanes gather only; at most 3, keeping the key within 32 bits).
int render_sub_bits(const Ctx* c, uint32_t size);
// Builds the grid (sort by bucket of (pos4 or pos3) through perm) into the
// context's sort buffers, avoiding the buffer `keep` (index into u[], or -1).
// Returns the buffer holding the sorted values (grid sorted_idx).  With a
// device count the n slots hold *count records followed by unused ones, whose
// keys sort behind every bucket.
uint32_t* build_grid(Ctx* c, const float4* pos4, const float* pos3, const uint32_t* perm, int64_t n,
                     const GridParams& g, int keep, const unsigned long long* count = nullptr,
                     uint32_t* off_out = nullptr, bool keys_ready = false);
// The photon records in origin order (photon index, then rank); with
// starts_done the per-photon record starts (origin_starts, which needs only
// the photon pass's per-photon counts) are already enqueued on s.
uint32_t* origin_order(Ctx* c, hipStream_t s, int64_t n, const GridParams& g, const unsigned long long* count,
                       bool starts_done = false);
void origin_starts(Ctx* c, hipStream_t s);
hipError_t ensure_sort_buffers(Ctx* c, int64_t n, uint32_t S);
hipError_t ensure_hit_sort(Ctx* c, int64_t n);
void sort_hitpoints(hipStream_t s, DBuf* hk, DBuf& hhist, const float4* hp, const float4* hw, int64_t cap,
                    const GridParams& g, double radius, float4* out_pos, float4* out_w);
void sort_hitpoints(Ctx* c, hipStream_t s, const float4* hp, const float4* hw, int64_t cap, const GridParams& g,
                    double radius);
void copy_ranges(const std::vector<uint32_t>& off, int32_t S, int32_t* cs, int32_t* ce);
EyePassArgs eye_args(Ctx* ctx, uint32_t pass);
// subsurface scattering (sss_host.hip)
double dipole_fdr(double eta);
uint32_t morton_bits(uint32_t v);
hipError_t subsurface_setup(Ctx* c, const float* tri_h /* may be null: read back */);
hipError_t subsurface_irradiance(Ctx* c, int64_t n_irr, double radius);
// the octree cut of the dipole sum (dipole_cut.hip): the preorder tree built
// on the host, its upload, its device view and the per-pass node values
struct CutHost {
    std::vector<float4> lo, hi, a;
    std::vector<uint32_t> esc, perm;
    std::vector<int32_t> depth;
};
void cut_build(const float4* pts, int64_t n, double area, CutHost* out);
hipError_t cut_upload(CutBufs* t, const CutHost& h);
CutDev cut_dev(const CutBufs& t, double max_error);
void cut_refresh(const CutBufs& t, hipStream_t s, const float* E, double max_error);
// no-allocation scratch (noalloc.hip): its device address on `device` (the
// current device), looked up once per device
hipError_t noalloc_setup(int device, void** addr);
// environment map (envmap_host.hip): the importance cells of Envmap::
// createImportanceMap (FP64, bit for bit), and the emission constants of a
// render with N photons per pass (area-light power total in, pick and disc
// constants out)
void envmap_importance(const float* rgb, int32_t w, int32_t h, double* pdf, double* cdf, int32_t* last);
double envmap_power(const Ctx* c);
// the Hachisuka method (hachisuka.hip): render-point state, one pass (an
// error message, empty on success), the FP64 image into device memory
hipError_t hachisuka_prepare(Ctx* c);
std::string hachisuka_pass(Ctx* c, pppm_pass_stats* st);
hipError_t hachisuka_image(Ctx* c, double* out_dev);
// learned emission proposal (proposal_host.hip)
uint64_t splitmix64(uint64_t* st);
std::vector<float> mlp_init_params(uint64_t seed);
hipError_t mlp_upload(Ctx* c, const float* params);
hipError_t ensure_mlp(Ctx* c);
hipError_t ensure_train(Ctx* c, int64_t B);
void mlp_step(Ctx* c, const float4* rec, int64_t n_valid, int64_t B, bool identity, double lr);
hipError_t proposal_tables_from_mlp(Ctx* c, int32_t side);
hipError_t proposal_retrain(Ctx* c, double* loss);

}  // namespace pppmi
