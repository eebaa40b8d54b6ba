// subsurface.hip -- the dipole diffusion sum of translucent surfaces on gfx950.
//
// Restates SubsurfaceIntegrator::irradiance (sources/subsurface_integrator.cc:
// 147-167, 234-238) with DipoleBSSRDF::operator() (sources/bssrdf.cc:92-104):
// for an eye path transmitted into a BSSRDF surface at x,
//   M_o(x) = sum_j R_d(|x - x_j|^2) E_j A_j,   L = (1 - F_dr) / pi * M_o,
// over all surface points j (the reference walks an octree cut of the same
// sum).  R_d per channel, with d+ = sqrt(d^2 + z+^2), d- = sqrt(d^2 + z-^2):
//   R_d = alpha' / (4 pi sigma_tr) (z+ (sigma_tr d+ + 1) e^{-sigma_tr d+} / d+^3
//                                   - z- (sigma_tr d- + 1) e^{-sigma_tr d-} / d-^3),
// clamped at 0, with the true exponential (the reference's Schraudolph
// approximation, sources/vector3d.cc:11-16, is not reproduced).
//
// N-body layout: 256 eye hits per block in registers, the points staged
// through LDS in tiles of 256 (position and irradiance), every lane walking
// the tile from the same address (broadcast reads).  Sums in FP32 with Kahan
// compensation per channel.
#include "internal.h"

namespace {

#define DIP_TILE 256

struct Kahan3 {
    float s[3] = {0.0f, 0.0f, 0.0f};
    float c[3] = {0.0f, 0.0f, 0.0f};
};

// one channel of R_d(d2) (without the alpha' / (4 pi sigma_tr) factor)
__device__ __forceinline__ float dipole_term(float d2, float zp, float zn, float tr) {
    const float ip = rsqrtf(d2 + zp * zp), in = rsqrtf(d2 + zn * zn);
    const float dp = (d2 + zp * zp) * ip, dn = (d2 + zn * zn) * in;
    const float tp = zp * (tr * dp + 1.0f) * __expf(-tr * dp) * (ip * ip * ip);
    const float tn = zn * (tr * dn + 1.0f) * __expf(-tr * dn) * (in * in * in);
    return tp - tn;
}

__global__ __launch_bounds__(256) void k_dipole(DipoleArgs a) {
    __shared__ float4 s_pos[DIP_TILE];
    __shared__ float s_e[3][DIP_TILE];
    const int64_t n = a.count_src ? (int64_t)a.count_src->sss_hits : a.n_hit;
    const int64_t base = (int64_t)blockIdx.x * 256;
    if (base >= n) return;  // the whole block: no barrier is left waiting
    const int64_t i = base + threadIdx.x;
    const bool active = i < n;
    float4 hp = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    DipoleDev D;
    if (active) {
        hp = a.hit_pos[i];
        D = a.dip[__float_as_int(hp.w)];
    } else {
        D = a.dip[0];
    }
    Kahan3 acc;
    for (int64_t t0 = 0; t0 < a.n_pts; t0 += DIP_TILE) {
        __syncthreads();
        const int64_t j = t0 + threadIdx.x;
        if (j < a.n_pts) {
            s_pos[threadIdx.x] = a.pts[j];
            s_e[0][threadIdx.x] = a.E[3 * j];
            s_e[1][threadIdx.x] = a.E[3 * j + 1];
            s_e[2][threadIdx.x] = a.E[3 * j + 2];
        } else {
            s_pos[threadIdx.x] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            s_e[0][threadIdx.x] = s_e[1][threadIdx.x] = s_e[2][threadIdx.x] = 0.0f;
        }
        __syncthreads();
        if (!active) continue;
        for (int k = 0; k < DIP_TILE; k++) {
            const float4 p = s_pos[k];
            const float dx = hp.x - p.x, dy = hp.y - p.y, dz = hp.z - p.z;
            const float d2 = dx * dx + dy * dy + dz * dz;
#pragma unroll
            for (int c = 0; c < 3; c++) {
                const float r = fmaxf(D.k[c] * dipole_term(d2, D.zp[c], D.zn[c], D.tr[c]), 0.0f);
                const float y = r * s_e[c][k] - acc.c[c];
                const float t = acc.s[c] + y;
                acc.c[c] = (t - acc.s[c]) - y;
                acc.s[c] = t;
            }
        }
    }
    if (!active) return;
    const float4 w = a.hit_w[i];
    const float f = D.out * a.area;
    const float L[3] = {w.x * f * acc.s[0], w.y * f * acc.s[1], w.z * f * acc.s[2]};
    const int64_t pix = (int64_t)__float_as_uint(w.w);
    if (a.img_sum) {
        for (int c = 0; c < 3; c++) kahan_add(a.img_sum, a.img_comp, 3 * pix + c, L[c]);
    } else {
        for (int c = 0; c < 3; c++) a.out_rgb[3 * i + c] = L[c];
    }
}

}  // namespace

void launch_dipole(hipStream_t s, const DipoleArgs& a) {
    if (a.n_hit <= 0 || a.n_pts <= 0) return;
    hipLaunchKernelGGL(k_dipole, dim3((unsigned)((a.n_hit + 255) / 256)), dim3(256), 0, s, a);
}
