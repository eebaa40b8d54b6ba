// proposal.hip -- learned photon-emission proposal on gfx950.
//
// No reference counterpart: it replaces the cosine emission of
// sources/sampler.cc:17-32 (called by ppm_probabilistic.cc:136-146 through
// the light contract of sources/envmap.cc:62-86) with the mixture of
// pppm_device.h, whose pdf is exact, so the estimator stays unbiased.
//
//  * Contribution records: after the gather of a pass, every emitted photon
//    with deposits that some hit point's gather accepted (inside the radius,
//    normals agreeing) gets a record (u, v, cell, c): its light-sample
//    coordinates, the direction cell it left in and the summed mean flux of
//    those deposits.  Records are written in
//    photon order (a prefix scan, no atomics) into a ring.
//  * MLP: Fourier encoding of (u, v) (32) -> 64 -> 64 -> 64 (ReLU) -> 1024
//    logits, BF16 weights and activations on gfx950's v_mfma_f32_16x16x32_bf16
//    with FP32 accumulation, the output layer split across the four waves of
//    a workgroup; training minimises the contribution-weighted cross
//    entropy of the record's cell (the KL divergence to the contribution
//    density); Adam in FP32.  Weight gradients are split-K partial sums
//    reduced in a fixed order, so training is deterministic.
//  * Tables: softmax of the logits at the anchors (the centres of a side x side
//    grid over the light) -> per-anchor inclusive CDF for the photon kernel.
#include "internal.h"

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ uint16_t to_bf16(float f) {
    uint32_t u = __float_as_uint(f);
    u += 0x7FFFu + ((u >> 16) & 1u);  // round to nearest even
    return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float from_bf16(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

// D += A B for one 16 x 16 x 32 tile.  Lane l = r + 16 g holds A[r][8g .. 8g+7],
// B[8g .. 8g+7][r] and D[4g + i][r] (i = 0..3).
__device__ __forceinline__ f32x4 mfma32(bf16x8 a, bf16x8 b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}
// eight BF16 values from a 16-byte aligned address
__device__ __forceinline__ bf16x8 ld8(const uint16_t* p) { return __builtin_bit_cast(bf16x8, *(const uint4*)p); }
__device__ __forceinline__ float bf16_sum8(bf16x8 v) {
    const uint4 u = __builtin_bit_cast(uint4, v);
    const uint32_t w[4] = {u.x, u.y, u.z, u.w};
    float s = 0.0f;
#pragma unroll
    for (int i = 0; i < 4; i++) s += __uint_as_float(w[i] << 16) + __uint_as_float(w[i] & 0xFFFF0000u);
    return s;
}

// four consecutive samples of feature row `row` of a feature-major [F][n] array
__device__ __forceinline__ void st4(uint16_t* base, int64_t n, int row, int64_t s, const uint16_t v[4]) {
    uint2 w;
    w.x = (uint32_t)v[0] | ((uint32_t)v[1] << 16);
    w.y = (uint32_t)v[2] | ((uint32_t)v[3] << 16);
    *(uint2*)(base + (int64_t)row * n + s) = w;
}

// Orders a wave's LDS writes before its later LDS reads (one wave's LDS
// operations are served in order; this keeps the compiler from moving them).
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

#define LDX 40  // LDS row strides in bf16: 16-byte aligned rows, padded against bank conflicts
#define LDH 72
#define LDG 40

// Fourier feature k of coordinate x: k = 2q (sin) or 2q + 1 (cos) of pi 2^q x,
// with the argument reduced exactly in revolutions
__device__ __forceinline__ void fourier(float x, int q, float* sn, float* cs) {
    const float t = x * (float)(1 << q) * 0.5f;  // revolutions, exact
    const float fr = t - floorf(t);
    sincosf(2.0f * PPPM_PI * fr, sn, cs);
}

// H[16 x 64] = relu(A[16 x K] W^T + b) for one wave (A, H in LDS; W [64][K] BF16)
template <int K>
__device__ __forceinline__ void dense_relu(const uint16_t* A, int lda, const uint16_t* W, const float* b,
                                           uint16_t* H, uint16_t* act, int64_t n, int act_row, int64_t s0) {
    const int l = threadIdx.x & 63, r = l & 15, g = l >> 4;
    bf16x8 av[K / 32];
#pragma unroll
    for (int kt = 0; kt < K / 32; kt++) av[kt] = ld8(A + r * lda + kt * 32 + 8 * g);
#pragma unroll
    for (int nt = 0; nt < MLP_HID / 16; nt++) {
        f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int kt = 0; kt < K / 32; kt++) acc = mfma32(av[kt], ld8(W + (nt * 16 + r) * K + kt * 32 + 8 * g), acc);
        const int col = nt * 16 + r;
        const float bc = b[col];
        uint16_t hv[4];
#pragma unroll
        for (int i = 0; i < 4; i++) {
            hv[i] = to_bf16(fmaxf(acc[i] + bc, 0.0f));
            H[(4 * g + i) * LDH + col] = hv[i];
        }
        if (act) st4(act, n, act_row + col, s0 + 4 * g, hv);
    }
}

// dA_out[16 x 64] = (dA_in[16 x 64] W) * [H > 0] for one wave (WT = W^T, [64][64])
__device__ __forceinline__ void dense_back(const uint16_t* Din, const uint16_t* WT, const uint16_t* H,
                                           uint16_t* Dout, uint16_t* act, int64_t n, int act_row, int64_t s0) {
    const int l = threadIdx.x & 63, r = l & 15, g = l >> 4;
    bf16x8 av[2];
#pragma unroll
    for (int kt = 0; kt < 2; kt++) av[kt] = ld8(Din + r * LDH + kt * 32 + 8 * g);
#pragma unroll
    for (int nt = 0; nt < 4; nt++) {
        f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int kt = 0; kt < 2; kt++) acc = mfma32(av[kt], ld8(WT + (nt * 16 + r) * 64 + kt * 32 + 8 * g), acc);
        const int col = nt * 16 + r;
        uint16_t dv[4];
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const float hv = from_bf16(H[(4 * g + i) * LDH + col]);
            dv[i] = to_bf16(hv > 0.0f ? acc[i] : 0.0f);
            if (Dout) Dout[(4 * g + i) * LDH + col] = dv[i];
        }
        st4(act, n, act_row + col, s0 + 4 * g, dv);
    }
}

// Forward (INFER: softmax probabilities) or forward + backward to the
// activation gradients (TRAIN) of 16 samples per workgroup of four waves.
// Wave 0 encodes the input and runs the hidden layers (and, for TRAIN, the
// backward through them); the 1024-way output layer is split by column, wave
// w taking columns [256 w, 256 w + 256): the softmax statistics combine the
// four waves' partial maxima and sums in wave order, and dH3 the four
// partial products in wave order, so the result does not depend on timing.
// A batch of B samples is B / 16 workgroups: four waves per SIMD at the
// training batch of 16384.
#define MLP_COLS_PER_WAVE (MLP_OUT / 4)
template <bool TRAIN>
__global__ __launch_bounds__(256) void k_mlp(MlpArgs a) {
    __shared__ __align__(16) uint16_t sX[16 * LDX];
    __shared__ __align__(16) uint16_t sH[3][16 * LDH];
    __shared__ __align__(16) uint16_t sD[2][16 * LDH];
    __shared__ __align__(16) uint16_t sG[4][16 * LDG];
    __shared__ float sW[16];
    __shared__ int32_t sT[16];
    __shared__ float sM[4][16], sS[4][16];
    __shared__ float sDH[4][16][MLP_HID];
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63, r = l & 15, g = l >> 4;
    const int64_t s0 = (int64_t)blockIdx.x * 16;
    const int64_t n = a.n;
    const float* P = a.params;
    const uint16_t* WB = a.wb;
    uint16_t* act = TRAIN ? a.act : nullptr;
    uint16_t* H1 = sH[0];
    uint16_t* H2 = sH[1];
    uint16_t* H3 = sH[2];

    if (w == 0) {
        // input encoding: lane (r, g) -> sample s0 + r, features 8g .. 8g + 7
        const int64_t s = s0 + r;
        float u = 0.0f, v = 0.0f;
        if (TRAIN) {
            const uint32_t j = a.idx ? a.idx[s] : (uint32_t)s;
            const float4 rc = a.rec[j];
            u = rc.x;
            v = rc.y;
            if (g == 0) {
                sW[r] = rc.w * a.inv_wsum[0];
                sT[r] = (int32_t)__float_as_uint(rc.z);
            }
        } else if (s < n) {
            const float2 t = a.uv[s];
            u = t.x;
            v = t.y;
        }
        const float x = g < 2 ? u : v;
#pragma unroll
        for (int q = 0; q < 4; q++) {
            float sn, cs;
            fourier(x, (g & 1) * 4 + q, &sn, &cs);
            sX[r * LDX + 8 * g + 2 * q] = to_bf16(sn);
            sX[r * LDX + 8 * g + 2 * q + 1] = to_bf16(cs);
        }
        wave_sync();
        if (TRAIN) {  // X0^T: lane -> feature l / 2, samples 8 (l & 1) .. + 7
            const int f = l >> 1, h = l & 1;
            uint16_t xv[8];
#pragma unroll
            for (int t = 0; t < 8; t++) xv[t] = sX[(8 * h + t) * LDX + f];
            uint4 pk;
            pk.x = (uint32_t)xv[0] | ((uint32_t)xv[1] << 16);
            pk.y = (uint32_t)xv[2] | ((uint32_t)xv[3] << 16);
            pk.z = (uint32_t)xv[4] | ((uint32_t)xv[5] << 16);
            pk.w = (uint32_t)xv[6] | ((uint32_t)xv[7] << 16);
            *(uint4*)(act + (int64_t)(ACT_X0 + f) * n + s0 + 8 * h) = pk;
        }
        dense_relu<MLP_IN>(sX, LDX, WB + MW1, P + OB1, H1, act, n, ACT_H1, s0);
        wave_sync();
        dense_relu<MLP_HID>(H1, LDH, WB + MW2, P + OB2, H2, act, n, ACT_H2, s0);
        wave_sync();
        dense_relu<MLP_HID>(H2, LDH, WB + MW3, P + OB3, H3, act, n, ACT_H3, s0);
    }
    __syncthreads();

    // output layer over this wave's columns, pass 1: per-row running max and
    // sum of exponentials
    bf16x8 h3[2];
#pragma unroll
    for (int kt = 0; kt < 2; kt++) h3[kt] = ld8(H3 + r * LDH + kt * 32 + 8 * g);
    const uint16_t* W4 = WB + MW4;
    const float* b4 = P + OB4;
    const int nt0 = w * (MLP_COLS_PER_WAVE / 16), nt1 = nt0 + MLP_COLS_PER_WAVE / 16;
    float M[4], S[4];
#pragma unroll
    for (int i = 0; i < 4; i++) {
        M[i] = -INFINITY;
        S[i] = 0.0f;
    }
    // four tiles per step, their weights and biases loaded before any is used
    for (int nb = nt0; nb < nt1; nb += 4) {
        bf16x8 wv[4][2];
        float bcv[4];
#pragma unroll
        for (int q = 0; q < 4; q++) {
#pragma unroll
            for (int kt = 0; kt < 2; kt++) wv[q][kt] = ld8(W4 + ((nb + q) * 16 + r) * 64 + kt * 32 + 8 * g);
            bcv[q] = b4[(nb + q) * 16 + r];
        }
#pragma unroll
        for (int q = 0; q < 4; q++) {
            f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
            for (int kt = 0; kt < 2; kt++) acc = mfma32(h3[kt], wv[q][kt], acc);
#pragma unroll
            for (int i = 0; i < 4; i++) {
                const float z = acc[i] + bcv[q];
                if (z > M[i]) {
                    S[i] = S[i] * __expf(M[i] - z) + 1.0f;
                    M[i] = z;
                } else {
                    S[i] += __expf(z - M[i]);
                }
            }
        }
    }
#pragma unroll
    for (int off = 1; off < 16; off <<= 1) {
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const float m2 = __shfl_xor(M[i], off), s2 = __shfl_xor(S[i], off);
            const float mm = fmaxf(M[i], m2);
            S[i] = S[i] * __expf(M[i] - mm) + s2 * __expf(m2 - mm);
            M[i] = mm;
        }
    }
    if (r == 0)
#pragma unroll
        for (int i = 0; i < 4; i++) {
            sM[w][4 * g + i] = M[i];
            sS[w][4 * g + i] = S[i];
        }
    __syncthreads();
    // the four waves' statistics of each row, combined in wave order
#pragma unroll
    for (int i = 0; i < 4; i++) {
        float mm = sM[0][4 * g + i], ss = sS[0][4 * g + i];
#pragma unroll
        for (int v = 1; v < 4; v++) {
            const float m2 = sM[v][4 * g + i], s2 = sS[v][4 * g + i];
            const float mx = fmaxf(mm, m2);
            ss = ss * __expf(mm - mx) + s2 * __expf(m2 - mx);
            mm = mx;
        }
        M[i] = mm;
        S[i] = ss;
    }
    float invS[4];
#pragma unroll
    for (int i = 0; i < 4; i++) invS[i] = 1.0f / S[i];

    if (!TRAIN) {
        for (int nb = nt0; nb < nt1; nb += 4) {
            bf16x8 wv[4][2];
            float bcv[4];
#pragma unroll
            for (int q = 0; q < 4; q++) {
#pragma unroll
                for (int kt = 0; kt < 2; kt++) wv[q][kt] = ld8(W4 + ((nb + q) * 16 + r) * 64 + kt * 32 + 8 * g);
                bcv[q] = b4[(nb + q) * 16 + r];
            }
#pragma unroll
            for (int q = 0; q < 4; q++) {
                f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
                for (int kt = 0; kt < 2; kt++) acc = mfma32(h3[kt], wv[q][kt], acc);
#pragma unroll
                for (int i = 0; i < 4; i++) {
                    const int64_t s = s0 + 4 * g + i;
                    if (s < n) a.prob[s * MLP_OUT + (nb + q) * 16 + r] = __expf(acc[i] + bcv[q] - M[i]) * invS[i];
                }
            }
        }
        return;
    }

    // pass 2: dZ = w (softmax - onehot) over this wave's columns, and its
    // share of dH3 = dZ W4
    const uint16_t* W4T = WB + MW4T;
    uint16_t* G = sG[w];
    f32x4 dh[4];
#pragma unroll
    for (int t = 0; t < 4; t++) dh[t] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
    float wt[4];
    int32_t tg[4];
#pragma unroll
    for (int i = 0; i < 4; i++) {
        wt[i] = sW[4 * g + i];
        tg[i] = sT[4 * g + i];
    }
    float lpart = 0.0f;
    for (int nt2 = nt0 / 2; nt2 < nt1 / 2; nt2++) {  // 32 output columns: one K step of dH3
        // this step's weights (both halves), biases and transposed weights,
        // loaded before any is used
        bf16x8 wv[2][2], wt4[4];
        float bcv[2];
#pragma unroll
        for (int half = 0; half < 2; half++) {
#pragma unroll
            for (int kt = 0; kt < 2; kt++)
                wv[half][kt] = ld8(W4 + ((2 * nt2 + half) * 16 + r) * 64 + kt * 32 + 8 * g);
            bcv[half] = b4[(2 * nt2 + half) * 16 + r];
        }
#pragma unroll
        for (int ht = 0; ht < 4; ht++) wt4[ht] = ld8(W4T + (ht * 16 + r) * MLP_OUT + nt2 * 32 + 8 * g);
#pragma unroll
        for (int half = 0; half < 2; half++) {
            const int nt = 2 * nt2 + half;
            f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
            for (int kt = 0; kt < 2; kt++) acc = mfma32(h3[kt], wv[half][kt], acc);
            const int col = nt * 16 + r;
            const float bc = bcv[half];
            uint16_t gv[4];
#pragma unroll
            for (int i = 0; i < 4; i++) {
                const float z = acc[i] + bc;
                const float p = __expf(z - M[i]) * invS[i];
                const bool hit = col == tg[i];
                if (hit) lpart += wt[i] * (M[i] + __logf(S[i]) - z);
                gv[i] = to_bf16(wt[i] * (p - (hit ? 1.0f : 0.0f)));
                G[(4 * g + i) * LDG + half * 16 + r] = gv[i];
            }
            st4(act, n, ACT_DZ + col, s0 + 4 * g, gv);
        }
        wave_sync();
        const bf16x8 ga = ld8(G + r * LDG + 8 * g);
#pragma unroll
        for (int ht = 0; ht < 4; ht++) dh[ht] = mfma32(ga, wt4[ht], dh[ht]);
        wave_sync();
    }
    if (a.loss) {
        for (int off = 32; off > 0; off >>= 1) lpart += __shfl_xor(lpart, off);
        if (l == 0) atomicAdd(a.loss, lpart);
    }
#pragma unroll
    for (int ht = 0; ht < 4; ht++)
#pragma unroll
        for (int i = 0; i < 4; i++) sDH[w][4 * g + i][ht * 16 + r] = dh[ht][i];
    __syncthreads();
    if (w != 0) return;
    // dA3 = (sum of the four waves' dH3, in wave order) * [H3 > 0]
    uint16_t* D0 = sD[0];
    uint16_t* D1 = sD[1];
#pragma unroll
    for (int ht = 0; ht < 4; ht++) {
        const int col = ht * 16 + r;
        uint16_t dv[4];
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const int row = 4 * g + i;
            const float d = ((sDH[0][row][col] + sDH[1][row][col]) + sDH[2][row][col]) + sDH[3][row][col];
            const float hv = from_bf16(H3[row * LDH + col]);
            dv[i] = to_bf16(hv > 0.0f ? d : 0.0f);
            D0[row * LDH + col] = dv[i];
        }
        st4(act, n, ACT_D3 + col, s0 + 4 * g, dv);
    }
    wave_sync();
    dense_back(D0, WB + MW3T, H2, D1, act, n, ACT_D2, s0);
    wave_sync();
    dense_back(D1, WB + MW2T, H1, nullptr, act, n, ACT_D1, s0);
}

// Split-K weight gradient of one layer: for K chunk ks,
//   part[ks][off + m N + c] = sum_k GT[m][k] HT[c][k],  part[ks][off + M N + m] = sum_k GT[m][k]
// One wave per (16-row tile, chunk).
template <int N>
__global__ __launch_bounds__(64) void k_mlp_wgrad(const uint16_t* GT, const uint16_t* HT, int64_t B, int M,
                                                  float* part, int64_t part_stride, int off) {
    const int mt = blockIdx.x, ks = blockIdx.y;
    const int l = threadIdx.x, r = l & 15, g = l >> 4;
    const int64_t chunk = B / MLP_SPLITS;
    const int64_t k0 = (int64_t)ks * chunk;
    f32x4 acc[N / 16];
#pragma unroll
    for (int t = 0; t < N / 16; t++) acc[t] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
    float bs = 0.0f;
    const uint16_t* arow = GT + (int64_t)(mt * 16 + r) * B;
    // two K steps per iteration, every load issued before the products; the
    // chunk (a multiple of 32) may end with one single step
    const int64_t k1 = k0 + chunk;
    for (int64_t k = k0; k + 64 <= k1; k += 64) {
        bf16x8 av[2], hv[2][N / 16];
#pragma unroll
        for (int u = 0; u < 2; u++) {
            av[u] = ld8(arow + k + 32 * u + 8 * g);
#pragma unroll
            for (int t = 0; t < N / 16; t++) hv[u][t] = ld8(HT + (int64_t)(t * 16 + r) * B + k + 32 * u + 8 * g);
        }
#pragma unroll
        for (int u = 0; u < 2; u++) {
            bs += bf16_sum8(av[u]);
#pragma unroll
            for (int t = 0; t < N / 16; t++) acc[t] = mfma32(av[u], hv[u][t], acc[t]);
        }
    }
    if ((chunk & 63) != 0) {
        const int64_t k = k1 - 32;
        const bf16x8 av = ld8(arow + k + 8 * g);
        bs += bf16_sum8(av);
#pragma unroll
        for (int t = 0; t < N / 16; t++) acc[t] = mfma32(av, ld8(HT + (int64_t)(t * 16 + r) * B + k + 8 * g), acc[t]);
    }
    float* Pp = part + (int64_t)ks * part_stride + off;
#pragma unroll
    for (int t = 0; t < N / 16; t++)
#pragma unroll
        for (int i = 0; i < 4; i++) Pp[(mt * 16 + 4 * g + i) * N + t * 16 + r] = acc[t][i];
    bs += __shfl_xor(bs, 16);
    bs += __shfl_xor(bs, 32);
    if (g == 0) Pp[M * N + mt * 16 + r] = bs;
}

// BF16 mirror entries of master parameter p (transposed copies for the
// backward products)
__device__ inline void mirror_write(uint16_t* wb, int p, float w) {
    const uint16_t h = to_bf16(w);
    if (p < OB1) {
        wb[MW1 + p] = h;
    } else if (p >= OW2 && p < OB2) {
        const int q = p - OW2;
        wb[MW2 + q] = h;
        wb[MW2T + (q % 64) * 64 + q / 64] = h;
    } else if (p >= OW3 && p < OB3) {
        const int q = p - OW3;
        wb[MW3 + q] = h;
        wb[MW3T + (q % 64) * 64 + q / 64] = h;
    } else if (p >= OW4 && p < OB4) {
        const int q = p - OW4;
        wb[MW4 + q] = h;
        wb[MW4T + (q % 64) * MLP_OUT + q / 64] = h;
    }
}

__global__ void k_mlp_mirror(const float* params, uint16_t* wb) {
    const int p = blockIdx.x * 256 + threadIdx.x;
    if (p < MLP_NPARAM) mirror_write(wb, p, params[p]);
}

// gradient = fixed-order sum of the split-K partials; optional Adam step
__global__ void k_mlp_adam(AdamArgs a) {
    const int p = blockIdx.x * 256 + threadIdx.x;
    if (p >= MLP_NPARAM) return;
    float gsum = 0.0f;
    for (int ks = 0; ks < MLP_SPLITS; ks++) gsum += a.part[(int64_t)ks * a.part_stride + p];
    if (a.grad_out) a.grad_out[p] = gsum;
    if (!a.apply) return;
    const float m = a.b1 * a.m[p] + (1.0f - a.b1) * gsum;
    const float v = a.b2 * a.v[p] + (1.0f - a.b2) * gsum * gsum;
    a.m[p] = m;
    a.v[p] = v;
    const float w = a.params[p] - a.lr * (m * a.c1) / (sqrtf(v * a.c2) + a.eps);
    a.params[p] = w;
    mirror_write(a.wb, p, w);
}

// Mini-batch: record indices (uniform over the ring, or the identity) and
// 1 / sum of their weights, summed in a fixed order
__global__ __launch_bounds__(1024) void k_mlp_prep(const float4* ring, int64_t n_valid, int64_t B, uint64_t seed,
                                                   uint32_t step, int identity, uint32_t* idx, float* inv_wsum) {
    __shared__ float sh[1024];
    const int t = threadIdx.x;
    const uint2 key = make_uint2((uint32_t)seed, (uint32_t)(seed >> 32));
    float acc = 0.0f;
    for (int64_t s = t; s < B; s += 1024) {
        uint32_t j;
        if (identity) {
            j = (uint32_t)s;
        } else {
            const uint4 c = make_uint4((uint32_t)s, step, 0u, KIND_TRAIN << 24);
            const uint4 x = philox4x32_10(c, key);
            j = (uint32_t)(((uint64_t)x.x * (uint64_t)n_valid) >> 32);
        }
        idx[s] = j;
        acc += ring[j].w;
    }
    sh[t] = acc;
    __syncthreads();
    for (int off = 512; off > 0; off >>= 1) {
        if (t < off) sh[t] += sh[t + off];
        __syncthreads();
    }
    if (t == 0) inv_wsum[0] = sh[0] > 0.0f ? 1.0f / sh[0] : 0.0f;
}

// Per-anchor inclusive CDF of a probability row, summed sequentially; the
// last entry is 1 (an all-zero row becomes uniform)
__global__ void k_prop_table(const float* prob, int32_t anchors, float* cdf) {
    const int a = blockIdx.x * 64 + threadIdx.x;
    if (a >= anchors) return;
    const float* p = prob + (int64_t)a * PROP_CELLS;
    float* c = cdf + (int64_t)a * PROP_CELLS;
    float s = 0.0f;
    for (int k = 0; k < PROP_CELLS; k++) {
        s += fmaxf(p[k], 0.0f);
        c[k] = s;
    }
    if (s > 0.0f) {
        for (int k = 0; k < PROP_CELLS - 1; k++) c[k] = c[k] / s;
    } else {
        for (int k = 0; k < PROP_CELLS - 1; k++) c[k] = (float)(k + 1) * (1.0f / PROP_CELLS);
    }
    c[PROP_CELLS - 1] = 1.0f;
}

__device__ __forceinline__ uint32_t photon_of(const ContribArgs& a, int64_t k) {
    return __float_as_uint(a.ph_flux[(int64_t)a.perm[k] * a.rec_stride].w) >> 6;
}

// At the first record of each photon (records in origin order): the summed
// mean flux of its deposits that some hit point's gather accepted; the scan
// input is 1 where that is positive
__global__ __launch_bounds__(256) void k_contrib(ContribArgs a) {
#pragma clang fp contract(off)  // the sum is restated on the CPU bit for bit
    const int64_t k = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (k >= a.n) return;
    const uint32_t ph = photon_of(a, k);
    const bool first = k == 0 || photon_of(a, k - 1) != ph;
    float c = 0.0f;
    if (first) {
        for (int64_t j = k; j < a.n; j++) {
            const float4 f = a.ph_flux[(int64_t)a.perm[j] * a.rec_stride];
            if ((__float_as_uint(f.w) >> 6) != ph) break;
            if (a.used[a.perm[j]]) c += ((f.x + f.y) + f.z) * (1.0f / 3.0f);
        }
    }
    a.seg_c[k] = c;
    a.scan[k] = c > 0.0f ? 1u : 0u;
}

// Records at their scanned positions (the last ring_cap of the pass kept)
__global__ __launch_bounds__(256) void k_contrib_write(ContribArgs a) {
    const int64_t k = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (k >= a.n) return;
    const int64_t total = (int64_t)a.scan[a.n - 1] + (a.seg_c[a.n - 1] > 0.0f ? 1 : 0);
    if (k == a.n - 1) a.counters->proposal_records = (unsigned long long)total;
    const float c = a.seg_c[k];
    if (!(c > 0.0f)) return;
    const int64_t pos = (int64_t)a.scan[k];
    if (pos < total - a.ring_cap) return;
    const uint32_t ph = photon_of(a, k);
    Rng rng;
    rng.init(a.seed, KIND_PHOTON, a.pass, ph, a.scene_id);
    const float pu = rng.at(1), pv = rng.at(2);
    const float r3 = rng.at(3), r4 = rng.at(4);
    float sel = 0.0f, j6 = 0.0f, j7 = 0.0f;
    if (a.prop.side) {
        sel = rng.at(PH_DIM_PROP_SEL);
        j6 = rng.at(PH_DIM_PROP_JIT);
        j7 = rng.at(PH_DIM_PROP_JIT + 1);
    }
    const PropPick pk = prop_pick(a.prop, pu, pv, r3, r4, sel, j6, j7);
    a.ring[(a.head + pos) % a.ring_cap] = make_float4(pu, pv, __uint_as_float(pk.cell), c);
}

}  // namespace

static inline unsigned grid1(int64_t n, int b) { return (unsigned)((n + b - 1) / b); }

void launch_mlp(hipStream_t s, const MlpArgs& a, bool train) {
    if (a.n <= 0) return;
    const unsigned blocks = grid1(a.n, 16);
    if (train) hipLaunchKernelGGL(k_mlp<true>, dim3(blocks), dim3(256), 0, s, a);
    else hipLaunchKernelGGL(k_mlp<false>, dim3(blocks), dim3(256), 0, s, a);
}

void launch_mlp_wgrad(hipStream_t s, const uint16_t* act, int64_t B, float* part) {
    const int64_t st = MLP_NPARAM;
    hipLaunchKernelGGL(k_mlp_wgrad<MLP_IN>, dim3(MLP_HID / 16, MLP_SPLITS), dim3(64), 0, s,
                       act + (int64_t)ACT_D1 * B, act + (int64_t)ACT_X0 * B, B, MLP_HID, part, st, OW1);
    hipLaunchKernelGGL(k_mlp_wgrad<MLP_HID>, dim3(MLP_HID / 16, MLP_SPLITS), dim3(64), 0, s,
                       act + (int64_t)ACT_D2 * B, act + (int64_t)ACT_H1 * B, B, MLP_HID, part, st, OW2);
    hipLaunchKernelGGL(k_mlp_wgrad<MLP_HID>, dim3(MLP_HID / 16, MLP_SPLITS), dim3(64), 0, s,
                       act + (int64_t)ACT_D3 * B, act + (int64_t)ACT_H2 * B, B, MLP_HID, part, st, OW3);
    hipLaunchKernelGGL(k_mlp_wgrad<MLP_HID>, dim3(MLP_OUT / 16, MLP_SPLITS), dim3(64), 0, s,
                       act + (int64_t)ACT_DZ * B, act + (int64_t)ACT_H3 * B, B, MLP_OUT, part, st, OW4);
}

void launch_mlp_adam(hipStream_t s, const AdamArgs& a) {
    hipLaunchKernelGGL(k_mlp_adam, dim3(grid1(MLP_NPARAM, 256)), dim3(256), 0, s, a);
}

void launch_mlp_mirror(hipStream_t s, const float* params, uint16_t* wb) {
    hipLaunchKernelGGL(k_mlp_mirror, dim3(grid1(MLP_NPARAM, 256)), dim3(256), 0, s, params, wb);
}

void launch_mlp_prep(hipStream_t s, const float4* ring, int64_t n_valid, int64_t B, uint64_t seed, uint32_t step,
                     int identity, uint32_t* idx, float* inv_wsum) {
    hipLaunchKernelGGL(k_mlp_prep, dim3(1), dim3(1024), 0, s, ring, n_valid, B, seed, step, identity, idx, inv_wsum);
}

void launch_prop_table(hipStream_t s, const float* prob, int32_t anchors, float* cdf) {
    if (anchors <= 0) return;
    hipLaunchKernelGGL(k_prop_table, dim3(grid1(anchors, 64)), dim3(64), 0, s, prob, anchors, cdf);
}

void launch_contrib(hipStream_t s, const ContribArgs& a, uint64_t* block_sums) {
    if (a.n <= 0) return;
    hipLaunchKernelGGL(k_contrib, dim3(grid1(a.n, 256)), dim3(256), 0, s, a);
    // the last flag and value are read by k_contrib_write through seg_c
    launch_exclusive_scan_u64(s, a.scan, a.n, block_sums);
    hipLaunchKernelGGL(k_contrib_write, dim3(grid1(a.n, 256)), dim3(256), 0, s, a);
}
