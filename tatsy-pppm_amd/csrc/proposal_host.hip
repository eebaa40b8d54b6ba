// proposal_host.hip -- the learned emission proposal, host side: MLP
// parameters, training steps, tables, and the pppm_proposal_* / pppm_mlp_*
// entry points.  No reference counterpart (it replaces the cosine emission of
// sources/sampler.cc:17-32).
#include "ctx.h"

using namespace pppmi;

namespace pppmi {

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


}  // namespace pppmi

extern "C" {

int pppm_proposal_enable(void* ctx_raw, const pppm_proposal_params* p) {
    CHECK_CTX(ctx);
    if (!ctx->rendering) return ctx->fail(PPPM_ERR_STATE, "call pppm_render_begin first");
    if (!p || !(p->lambda > 0.0 && p->lambda <= 1.0) || !(p->learning_rate >= 0.0) || p->anchors < 1 ||
        p->anchors > 32 || p->train_interval < 1 || p->train_steps < 1 || p->batch < 256 || p->batch % 256 != 0 ||
        p->batch > 65536 || p->ring_capacity < 1 || p->ring_capacity > (1LL << 28))
        return ctx->fail(PPPM_ERR_INVALID, "bad proposal parameters");
    if (ctx->h_lights.size() != 1 || ctx->has_env || ctx->rp.method != PPPM_METHOD_PROBABILISTIC)
        return ctx->fail(PPPM_ERR_INVALID, "the learned proposal needs the probabilistic method, exactly one area "
                                           "light and no environment map");
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
    if (ctx->h_lights.size() != 1 || ctx->has_env || ctx->rp.method != PPPM_METHOD_PROBABILISTIC)
        return ctx->fail(PPPM_ERR_INVALID, "the learned proposal needs the probabilistic method, exactly one area "
                                           "light and no environment map");
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
