// envmap_host.hip -- the environment-map emitter, host side (row f1):
// Envmap's importance map (sources/envmap.cc:88-163), the per-render
// emission constants, and pppm_envmap_set / pppm_envmap_importance.
//
// The importance map is the reference's, bit for bit: 64 x 64 cells over
// equal-area coordinates (u -> phi = 2 pi u, v -> y = 1 - 2 v), each weighted
// by the squared luminance of the mean of (W/64) x (H/64) texel lookups,
// normalised by (total + 1e-12), with running sums in cell order.  It is
// built once per map on the host in FP64 (the sums are sequential); the
// device draws from it per photon (render.hip).
#include "ctx.h"

namespace {

constexpr double kPi = 3.14159265358979323846;

// Envmap::sampleFromDir (envmap.cc:47-60) on the host image
void lookup(const float* rgb, int32_t w, int32_t h, double dx, double dy, double dz, double out[3]) {
    const double theta = std::acos(dy);
    double phi = std::atan2(dz, dx);
    if (phi < 0.0) phi += 2.0 * kPi;
    int ix = (int)(phi / (2.0 * kPi) * w), iy = (int)(theta / kPi * h);
    ix = std::min(std::max(ix, 0), w - 1);
    iy = std::min(std::max(iy, 0), h - 1);
    const float* p = rgb + 3 * ((size_t)iy * w + ix);
    out[0] = p[0];
    out[1] = p[1];
    out[2] = p[2];
}

}  // namespace

namespace pppmi {

void envmap_importance(const float* rgb, int32_t w, int32_t h, double* pdf, double* cdf, int32_t* last) {
    const int sx_n = w / ENV_CELLS, sy_n = h / ENV_CELLS;
    const double inv = 1.0 / (sx_n * sy_n);
    double total = 0.0;
    for (int iy = 0; iy < ENV_CELLS; iy++)
        for (int ix = 0; ix < ENV_CELLS; ix++) {
            double acc[3] = {0.0, 0.0, 0.0};
            for (int sy = 0; sy < sy_n; sy++)
                for (int sx = 0; sx < sx_n; sx++) {
                    const double u = (ix + (double)sx / sx_n) / ENV_CELLS;
                    const double v = (iy + (double)sy / sy_n) / ENV_CELLS;
                    const double phi = u * 2.0 * kPi;
                    const double y = (1.0 - v) * 2.0 - 1.0;
                    double c[3];
                    lookup(rgb, w, h, std::sqrt(1.0 - y * y) * std::cos(phi), y, std::sqrt(1.0 - y * y) * std::sin(phi),
                           c);
                    for (int k = 0; k < 3; k++) acc[k] += c[k] * inv;
                }
            const double l = acc[0] * 0.2126 + acc[1] * 0.7152 + acc[2] * 0.0722;  // luminance, vector3d.h:108-110
            pdf[iy * ENV_CELLS + ix] = l * l;
            total += l * l;
        }
    const double norm = total + 1.0e-12;
    pdf[0] /= norm;
    cdf[0] = pdf[0];
    for (int i = 1; i < ENV_CELLS * ENV_CELLS; i++) {
        pdf[i] /= norm;
        cdf[i] = pdf[i] + cdf[i - 1];
    }
    *last = 0;
    for (int i = 0; i < ENV_CELLS * ENV_CELLS; i++)
        if (pdf[i] > 0.0) *last = i;
}

// Selection power of the environment: pi R^2 times the integral over the
// sphere of the mean channel, texel by texel (texel solid angle
// (2 pi / w)(cos t0 - cos t1)); R is the radius of the triangles' bounding
// sphere (the emission disc).
double envmap_power(const Ctx* c) {
    const double dx = c->bmax[0] - c->bmin[0], dy = c->bmax[1] - c->bmin[1], dz = c->bmax[2] - c->bmin[2];
    const double R = 0.5 * std::sqrt(dx * dx + dy * dy + dz * dz);
    double pw = 0.0;
    for (int iy = 0; iy < c->env_h; iy++) {
        const double dom = (2.0 * kPi / c->env_w) * (std::cos(kPi * iy / c->env_h) - std::cos(kPi * (iy + 1) / c->env_h));
        double row = 0.0;
        for (int ix = 0; ix < c->env_w; ix++) {
            const float* p = &c->h_env[3 * ((size_t)iy * c->env_w + ix)];
            row += ((double)p[0] + (double)p[1] + (double)p[2]) / 3.0;
        }
        pw += row * dom;
    }
    return kPi * R * R * pw;
}

}  // namespace pppmi

using namespace pppmi;

extern "C" {

int pppm_envmap_importance(const float* rgb, int32_t width, int32_t height, double* pdf, double* cdf) {
    if (!rgb || !pdf || !cdf || width < ENV_CELLS || height < ENV_CELLS) return PPPM_ERR_INVALID;
    int32_t last;
    envmap_importance(rgb, width, height, pdf, cdf, &last);
    return PPPM_OK;
}

int pppm_envmap_set(void* ctx_raw, const float* rgb, int32_t width, int32_t height) {
    CHECK_CTX(ctx);
    if (!rgb) {
        ctx->has_env = false;
        ctx->h_env.clear();
        return PPPM_OK;
    }
    if (width < ENV_CELLS || height < ENV_CELLS || (int64_t)width * height > (1LL << 28))
        return ctx->fail(PPPM_ERR_INVALID, "environment maps are 64 x 64 texels or more");
    const size_t n = (size_t)width * height;
    for (size_t i = 0; i < 3 * n; i++)
        if (!(rgb[i] >= 0.0f) || !std::isfinite(rgb[i]))
            return ctx->fail(PPPM_ERR_INVALID, "environment texels must be finite and non-negative");
    (void)hipSetDevice(ctx->device);
    ctx->h_env.assign(rgb, rgb + 3 * n);
    ctx->h_env_pdf.assign(ENV_CELLS * ENV_CELLS, 0.0);
    ctx->h_env_cdf.assign(ENV_CELLS * ENV_CELLS, 0.0);
    envmap_importance(rgb, width, height, ctx->h_env_pdf.data(), ctx->h_env_cdf.data(), &ctx->env_last);
    std::vector<float4> tex(n);
    for (size_t i = 0; i < n; i++) tex[i] = make_float4(rgb[3 * i], rgb[3 * i + 1], rgb[3 * i + 2], 0.0f);
    TRY_HIP(ctx, ctx->env_tex.ensure(sizeof(float4) * n));
    TRY_HIP(ctx, ctx->env_pdf.ensure(sizeof(double) * ENV_CELLS * ENV_CELLS));
    TRY_HIP(ctx, ctx->env_cdf.ensure(sizeof(double) * ENV_CELLS * ENV_CELLS));
    TRY_HIP(ctx, hipMemcpy(ctx->env_tex.p, tex.data(), sizeof(float4) * n, hipMemcpyHostToDevice));
    TRY_HIP(ctx, hipMemcpy(ctx->env_pdf.p, ctx->h_env_pdf.data(), sizeof(double) * ENV_CELLS * ENV_CELLS,
                           hipMemcpyHostToDevice));
    TRY_HIP(ctx, hipMemcpy(ctx->env_cdf.p, ctx->h_env_cdf.data(), sizeof(double) * ENV_CELLS * ENV_CELLS,
                           hipMemcpyHostToDevice));
    ctx->env_w = width;
    ctx->env_h = height;
    ctx->has_env = true;
    ctx->rendering = false;  // a running render keeps its emitters: begin a new one
    return PPPM_OK;
}

}  // extern "C"
