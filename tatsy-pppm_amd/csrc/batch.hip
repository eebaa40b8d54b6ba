// batch.hip -- batched scenes: many renders in one hipGraph-captured pass
// loop (pppm_batch_*).
#include "ctx.h"

using namespace pppmi;

// ---------------------------------------------------------------------------
// Batched scenes: many renders in one hipGraph-captured pass loop.  No
// reference counterpart (the reference renders one scene per process,
// sources/main/appmain.cc); each scene's pass is the same computation as
// pppm_render_pass with a fixed hash size, enqueued without host
// synchronisation so that one capture holds a pass of every scene.
// ---------------------------------------------------------------------------
namespace {

using namespace pppmi;

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
void enqueue_pass(Ctx* ctx, int photon_grid, int trace_keep) {
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
    pa.trace_keep = trace_keep;
    launch_photon_pass(s, pa, photon_grid);
    launch_capacity_check(s, pd, dc, cap);
    GridParams g;
    for (int d = 0; d < 3; d++) g.bmin[d] = ctx->grid_bmin[d];
    g.scale = 0.0;
    g.size = (uint32_t)ctx->rp.hash_size;
    g.sub_bits = render_sub_bits(ctx, g.size);
    g.pd = pd;
    const int keep = 4;
    const uint32_t* perm1 = origin_order(ctx, s, cap, g, &dc->photons_stored);
    const uint32_t* sorted = build_grid(ctx, nullptr, nullptr, perm1, cap, g, keep,
                                        &dc->photons_stored, nullptr, true);
    launch_gather_layout(s, ctx->ph_rec.as<float4>(), ctx->ph_rec.as<float4>() + 1, nullptr, sorted, cap,
                         ctx->ph_a.as<float4>(), ctx->ph_b.as<float4>(),
                         gather_compact(ctx->rp.gather_mode) ? nullptr : ctx->ph_c.as<float4>(),
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
    int trace_keep = PHOTON_TRACE_KEEP;  // photon traversal's lanes in flight before regeneration (autotuned)
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
    if (params->method != PPPM_METHOD_PROBABILISTIC)
        return b->fail(PPPM_ERR_INVALID, "batched scenes render with the probabilistic method");
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
        if (tune_gather) modes = {GATHER_LANES, GATHER_BLOCKS, GATHER_TILED, GATHER_WAVE};
        else modes = {base.gather_mode};
        std::vector<int> grids;
        const int cu = c->cu > 0 ? c->cu : 256;
        if (tune_grid) grids = {cu, cu * 2, cu * 4, cu * 8};
        else grids = {sc.photon_grid};
        double best = 1e30;
        int best_mode = base.gather_mode, best_grid = sc.photon_grid, best_keep = sc.trace_keep;
        // one timed pass per candidate after a warm-up pass; returns false on error
        auto time_passes = [&](int mode, const std::vector<std::pair<int, int>>& shapes, bool& ok) {
            ok = false;
            pppm_render_params p = base;
            p.gather_mode = mode;
            p.passes = (int32_t)shapes.size() + 1;
            if (pppm_render_begin(c, &p) != PPPM_OK) return;
            if (prepare_enqueued(c) != hipSuccess) return;
            enqueue_pass(c, shapes[0].first, shapes[0].second);
            for (const auto& sh : shapes) {
                (void)hipEventRecord(c->ev[0], c->stream);
                enqueue_pass(c, sh.first, sh.second);
                (void)hipEventRecord(c->ev[1], c->stream);
                if (hipEventSynchronize(c->ev[1]) != hipSuccess) return;
                float ms = 0.0f;
                (void)hipEventElapsedTime(&ms, c->ev[0], c->ev[1]);
                if (ms < best) {
                    best = ms;
                    best_mode = mode;
                    best_grid = sh.first;
                    best_keep = sh.second;
                }
            }
            ok = true;
        };
        // gather kernel x photon grid, then the traversal's lanes in flight
        // at the best of those (coordinate descent)
        for (int m : modes) {
            std::vector<std::pair<int, int>> shapes;
            for (int gsz : grids) shapes.push_back({gsz, sc.trace_keep});
            bool ok;
            time_passes(m, shapes, ok);
            if (!ok) return b->fail(c->err.empty() ? PPPM_ERR_HIP : PPPM_ERR_STATE, "tuning pass: " + c->err);
        }
        if (tune_grid) {
            std::vector<std::pair<int, int>> shapes;
            for (int k : {8, 24, 32, 48}) shapes.push_back({best_grid, k});
            bool ok;
            time_passes(best_mode, shapes, ok);
            if (!ok) return b->fail(c->err.empty() ? PPPM_ERR_HIP : PPPM_ERR_STATE, "tuning pass: " + c->err);
        }
        pppm_render_params p = base;
        p.gather_mode = best_mode;
        int rc = pppm_render_begin(c, &p);
        if (rc != PPPM_OK) return b->fail(rc, c->err);
        if (prepare_enqueued(c) != hipSuccess) return b->fail(PPPM_ERR_HIP, "tuning reset");
        sc.photon_grid = best_grid;
        sc.trace_keep = best_keep;
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
            enqueue_pass(sc.ctx, sc.photon_grid, sc.trace_keep);
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

int pppm_batch_scene_trace_keep(void* batch, int32_t index, int32_t* trace_keep) {
    Batch* b = reinterpret_cast<Batch*>(batch);
    if (!b) return PPPM_ERR_INVALID;
    if (index < 0 || index >= (int32_t)b->scenes.size() || !trace_keep) return b->fail(PPPM_ERR_INVALID, "scene index");
    *trace_keep = b->scenes[(size_t)index].trace_keep;
    return PPPM_OK;
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

