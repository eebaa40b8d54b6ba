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
// the tile from the same address (broadcast reads).  Sums in FP32: a plain
// partial sum per tile, Kahan-compensated across tiles, per channel.
#include "internal.h"

namespace {

#define DIP_TILE 256

struct Kahan3 {
    float s[3] = {0.0f, 0.0f, 0.0f};
    float c[3] = {0.0f, 0.0f, 0.0f};
};

// Per-hit constants of one channel: z+^2, z-^2, sigma_tr log2(e) for exp2,
// z sigma_tr and z for the (sigma_tr d + 1) z factors, alpha' / (4 pi sigma_tr)
struct DipoleLane {
    float zp2[3], zn2[3], trl[3], ztp[3], zp[3], ztn[3], zn[3], k[3];
};

__device__ __forceinline__ DipoleLane dipole_lane(const DipoleDev& D) {
    DipoleLane L;
    const float log2e = 1.4426950408889634f;
#pragma unroll
    for (int c = 0; c < 3; c++) {
        L.zp2[c] = D.zp[c] * D.zp[c];
        L.zn2[c] = D.zn[c] * D.zn[c];
        L.trl[c] = D.tr[c] * log2e;
        L.ztp[c] = D.zp[c] * D.tr[c];
        L.zp[c] = D.zp[c];
        L.ztn[c] = D.zn[c] * D.tr[c];
        L.zn[c] = D.zn[c];
        L.k[c] = D.k[c];
    }
    return L;
}

// z (sigma_tr d + 1) e^{-sigma_tr d} / d^3 with d^2 = s (s >= z^2 > 0, so the
// raw reciprocal square root and exp2 need no denormal handling)
__device__ __forceinline__ float dipole_source(float s, float trl, float zt, float z) {
    const float id = __builtin_amdgcn_rsqf(s);
    const float d = s * id;
    const float e = __builtin_amdgcn_exp2f(-trl * d);
    return __builtin_fmaf(zt, d, z) * e * (id * id * id);
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
    const DipoleLane L = dipole_lane(D);
    // per tile a plain FP32 partial sum of 256 terms, compensated (Kahan)
    // across tiles
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
        float part[3] = {0.0f, 0.0f, 0.0f};
        for (int k = 0; k < DIP_TILE; k++) {
            const float4 p = s_pos[k];
            const float dx = hp.x - p.x, dy = hp.y - p.y, dz = hp.z - p.z;
            const float d2 = __builtin_fmaf(dx, dx, __builtin_fmaf(dy, dy, dz * dz));
#pragma unroll
            for (int c = 0; c < 3; c++) {
                const float r = dipole_source(d2 + L.zp2[c], L.trl[c], L.ztp[c], L.zp[c]) -
                                dipole_source(d2 + L.zn2[c], L.trl[c], L.ztn[c], L.zn[c]);
                part[c] = __builtin_fmaf(fmaxf(L.k[c] * r, 0.0f), s_e[c][k], part[c]);
            }
        }
#pragma unroll
        for (int c = 0; c < 3; c++) {
            const float y = part[c] - acc.c[c];
            const float t = acc.s[c] + y;
            acc.c[c] = (t - acc.s[c]) - y;
            acc.s[c] = t;
        }
    }
    if (!active) return;
    const float4 w = a.hit_w[i];
    const float f = D.out * a.area;
    const float Lo[3] = {w.x * f * acc.s[0], w.y * f * acc.s[1], w.z * f * acc.s[2]};
    const int64_t pix = (int64_t)__float_as_uint(w.w);
    if (a.img_sum) {
        for (int c = 0; c < 3; c++) kahan_add(a.img_sum, a.img_comp, 3 * pix + c, Lo[c]);
    } else {
        for (int c = 0; c < 3; c++) a.out_rgb[3 * i + c] = Lo[c];
    }
}

// Issue-rate probes: eight independent chains per lane so latency is hidden;
// one store per thread
#define PROBE_ITERS 4096
__global__ __launch_bounds__(256) void k_probe_fma(float* out, float a, float b) {
    float x[8];
#pragma unroll
    for (int k = 0; k < 8; k++) x[k] = (float)threadIdx.x * 1e-3f + (float)k;
    for (int i = 0; i < PROBE_ITERS; i++) {
#pragma unroll
        for (int k = 0; k < 8; k++) x[k] = __builtin_fmaf(x[k], a, b);
    }
    float s = 0.0f;
#pragma unroll
    for (int k = 0; k < 8; k++) s += x[k];
    out[(int64_t)blockIdx.x * 256 + threadIdx.x] = s;
}
__global__ __launch_bounds__(256) void k_probe_exp(float* out, float a, float b) {
    float x[8];
#pragma unroll
    for (int k = 0; k < 8; k++) x[k] = (float)threadIdx.x * 1e-6f + (float)k * 1e-3f;
    for (int i = 0; i < PROBE_ITERS; i++) {
#pragma unroll
        for (int k = 0; k < 8; k++) x[k] = __builtin_amdgcn_exp2f(x[k]) * a + b;
    }
    float s = 0.0f;
#pragma unroll
    for (int k = 0; k < 8; k++) s += x[k];
    out[(int64_t)blockIdx.x * 256 + threadIdx.x] = s;
}

}  // namespace

// wave instructions per launch: 8 chains x PROBE_ITERS per wave (FMA); the
// exp probe issues one v_exp_f32 and one v_fma_f32 per chain step
void launch_issue_probe(hipStream_t s, float* out, int blocks, int which) {
    if (which == 0) hipLaunchKernelGGL(k_probe_fma, dim3((unsigned)blocks), dim3(256), 0, s, out, 0.999f, 0.001f);
    else hipLaunchKernelGGL(k_probe_exp, dim3((unsigned)blocks), dim3(256), 0, s, out, 0.5f, 0.001f);
}
int64_t issue_probe_wave_instr(int which) { return (which == 0 ? 8 : 16) * (int64_t)PROBE_ITERS; }

void launch_dipole(hipStream_t s, const DipoleArgs& a) {
    if (a.n_hit <= 0 || a.n_pts <= 0) return;
    hipLaunchKernelGGL(k_dipole, dim3((unsigned)((a.n_hit + 255) / 256)), dim3(256), 0, s, a);
}
