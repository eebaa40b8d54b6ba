// pass.hip -- the pass loop of ProgressivePhotonMappingProb::render
// (sources/ppm_probabilistic.cc:24-121) split into pppm_render_begin /
// pppm_render_pass / pppm_render_passes: photon pass, origin order, hash
// grid, eye pass, hit-point sort and gather, single or pipelined on four
// streams.
#include "ctx.h"

using namespace pppmi;

namespace pppmi {

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

}  // namespace pppmi

extern "C" {

int pppm_render_begin(void* ctx_raw, const pppm_render_params* p) {
    CHECK_CTX(ctx);
    if (!ctx->has_bvh) return ctx->fail(PPPM_ERR_STATE, "no BVH built");
    if (!ctx->has_cam) return ctx->fail(PPPM_ERR_STATE, "no camera set");
    if (!p || p->photons <= 0 || p->passes <= 0 || !(p->alpha > 0.0 && p->alpha < 1.0) ||
        !(p->radius1 > 0.0) || p->max_bounces <= 0 || p->max_bounces > 64 || p->hash_size < 0 ||
        (p->gather_mode < PPPM_GATHER_AUTO || p->gather_mode > PPPM_GATHER_BLOCKS) ||
        (p->method != PPPM_METHOD_PROBABILISTIC && p->method != PPPM_METHOD_HACHISUKA))
        return ctx->fail(PPPM_ERR_INVALID, "bad render parameters");
    if (p->method == PPPM_METHOD_HACHISUKA && ctx->n_dips > 0)
        return ctx->fail(PPPM_ERR_INVALID, "the Hachisuka method takes no BSSRDF materials");
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
    // the environment map competes for photons by its power through the disc
    const double area_total = total;
    const double env_power = ctx->has_env ? envmap_power(ctx) : 0.0;
    total += env_power;
    if (!(total > 0.0)) return ctx->fail(PPPM_ERR_INVALID, "lights emit no power");
    ctx->pick_total = total;
    ctx->pick_area = area_total;
    if (ctx->has_env) {
        const double dx = ctx->bmax[0] - ctx->bmin[0], dy = ctx->bmax[1] - ctx->bmin[1], dz = ctx->bmax[2] - ctx->bmin[2];
        const double R = 0.5 * std::sqrt(dx * dx + dy * dy + dz * dz);
        ctx->env_k = M_PI * R * R * (4.0 * M_PI / (ENV_CELLS * ENV_CELLS)) / ((double)p->photons * (env_power / total));
    }
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
        o.cdf_hi = (l == nl - 1 && !ctx->has_env) ? 2.0f : (float)(cum / total);
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
    if (p->method == PPPM_METHOD_HACHISUKA) {
        TRY_HIP(ctx, hachisuka_prepare(ctx));
        TRY_HIP(ctx, hipStreamSynchronize(ctx->stream));
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
    if (ctx->rp.method == PPPM_METHOD_HACHISUKA) {
        const std::string e = hachisuka_pass(ctx, st);
        if (!e.empty()) return ctx->fail(e.find("photon buffer") == 0 ? PPPM_ERR_CAPACITY : PPPM_ERR_HIP, e);
        ctx->passes_done++;
        return PPPM_OK;
    }
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
    g.sub_bits = render_sub_bits(ctx, S);
    TRY_HIP(ctx, ctx->off.ensure(sizeof(uint32_t) * ((size_t)S + 1)));
    const int keep = 4;
    uint32_t* perm1 = origin_order(ctx, s, n, g, nullptr);
    TRY_HIP(ctx, hipEventRecord(ctx->ev[2], s));
    uint32_t* sorted = build_grid(ctx, nullptr, nullptr, perm1, n, g, keep, nullptr, nullptr, true);
    launch_gather_layout(s, ctx->ph_rec.as<float4>(), ctx->ph_rec.as<float4>() + 1, nullptr, sorted, n,
                         ctx->ph_a.as<float4>(), ctx->ph_b.as<float4>(),
                         gather_compact(ctx->rp.gather_mode) ? nullptr : ctx->ph_c.as<float4>(),
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
    const bool mark_in_gather = ctx->prop.enabled && n > 0 && gather_compact(ga.mode);
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
        if (ctx->sss_cut > 0.0) {
            cut_refresh(ctx->cut, s, ctx->sss_E.as<float>(), ctx->sss_cut);
            launch_dipole_cut(s, da, cut_dev(ctx->cut, ctx->sss_cut));
        } else {
            launch_dipole(s, da);
        }
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
        st->gather_staged = (int64_t)hc.gather_staged;
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
    if (ctx->prop.enabled || ctx->prop.active || ctx->sss_present || ctx->rp.method == PPPM_METHOD_HACHISUKA) {
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
    if (ctx->pinned_count < count) {
        (void)hipHostFree(ctx->pinned_counters);
        ctx->pinned_counters = nullptr;
        ctx->pinned_count = 0;
        TRY_HIP(ctx, hipHostMalloc((void**)&ctx->pinned_counters, sizeof(Counters) * (size_t)count,
                                   hipHostMallocDefault));
        ctx->pinned_count = count;
    }
    hipStream_t A = ctx->stream, Cs = ctx->stream3;
    hipStream_t Bq[2] = {ctx->stream2, ctx->stream4};
    DBuf* hkq[2] = {ctx->hk, ctx->hk_alt};
    DBuf* hhq[2] = {&ctx->hhist, &ctx->hhist_alt};
    enum { ZERO, P0, P1, O0, ODONE, E0, E1, HDONE, G0, GEND, GDONE };
    const int NEV = 6;  // event ring: pass k records into pev[k % NEV]
    const int eye_grid = ctx->cu > 0 ? ctx->cu * EYE_BLOCKS_PER_CU : 2048;
    const int photon_grid = ctx->cu > 0 ? ctx->cu * PHOTON_BLOCKS_PER_CU : 2048;
    const int32_t first = ctx->passes_done + 1;  // pass index of k = 0
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
        // the record starts per photon need only the photon kernel's counts;
        // the stored count goes to pinned memory, so the copy does not hold
        // the host, which enqueues the eye side of pass k + 2 while the
        // photons run (its event set was last used by pass k - 4, whose times
        // were read in the last iteration)
        origin_starts(ctx, A);
        Counters& hc = ctx->pinned_counters[k];
        TRY_HIP(ctx, hipMemcpyAsync(&hc, dc, sizeof(Counters), hipMemcpyDeviceToHost, A));
        if (k >= 3) read_times(k - 3);
        if (k + 2 < count) TRY_HIP(ctx, enqueue_eye(k + 2));
        // the stored count sizes the sorts; meanwhile the eye passes and the
        // previous gather run
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
        g.sub_bits = render_sub_bits(ctx, Sz);
        const int keep = 4;
        uint32_t* perm1 = origin_order(ctx, A, n, g, nullptr, true);
        ctx->stream = A;
        uint32_t* sorted = build_grid(ctx, nullptr, nullptr, perm1, n, g, keep, nullptr,
                                      S.off->as<uint32_t>(), true);
        ctx->stream = keep_stream;
        launch_gather_layout(A, ctx->ph_rec.as<float4>(), ctx->ph_rec.as<float4>() + 1, nullptr, sorted, n,
                             S.ph_a->as<float4>(), S.ph_b->as<float4>(),
                             gather_compact(ctx->rp.gather_mode) ? nullptr : S.ph_c->as<float4>(),
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
            st.gather_staged = (int64_t)c.gather_staged;
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

}  // extern "C"
