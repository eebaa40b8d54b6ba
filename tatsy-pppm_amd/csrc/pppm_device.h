// pppm_device.h -- device-side types and helpers shared by the gfx950 kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/pppm.h"

#define PPPM_PI 3.14159265358979323846f
#define PPPM_INV_PI 0.31830988618379067154f

// ---------------------------------------------------------------------------
// float3 arithmetic (explicit, so that FP contraction is visible at each site)
// ---------------------------------------------------------------------------
struct v3f {
    float x, y, z;
};
__host__ __device__ inline v3f mk(float x, float y, float z) { return v3f{x, y, z}; }
__host__ __device__ inline v3f operator+(v3f a, v3f b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
__host__ __device__ inline v3f operator-(v3f a, v3f b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
__host__ __device__ inline v3f operator*(v3f a, v3f b) { return mk(a.x * b.x, a.y * b.y, a.z * b.z); }
__host__ __device__ inline v3f operator*(v3f a, float s) { return mk(a.x * s, a.y * s, a.z * s); }
__host__ __device__ inline v3f operator-(v3f a) { return mk(-a.x, -a.y, -a.z); }
__host__ __device__ inline float dot(v3f a, v3f b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__host__ __device__ inline v3f cross(v3f u, v3f v) {
    return mk(u.y * v.z - u.z * v.y, u.z * v.x - u.x * v.z, u.x * v.y - u.y * v.x);
}
__host__ __device__ inline float vmax3(v3f a) { return fmaxf(a.x, fmaxf(a.y, a.z)); }
__device__ inline v3f normalize(v3f a) {
    float n2 = dot(a, a);
    return n2 > 0.0f ? a * rsqrtf(n2) : a;
}
__device__ inline v3f xyz(float4 a) { return mk(a.x, a.y, a.z); }
__device__ inline float get(v3f a, int k) { return k == 0 ? a.x : (k == 1 ? a.y : a.z); }

// ---------------------------------------------------------------------------
// Philox-4x32-10 (Salmon et al. 2011) counter-based RNG.
// key = seed, counter = (index, scene, pass, kind << 24 | block); one block
// gives four 24-bit uniforms.  Identical to the CPU oracle's stream.
// ---------------------------------------------------------------------------
__device__ inline uint4 philox4x32_10(uint4 c, uint2 k) {
#pragma unroll
    for (int r = 0; r < 10; r++) {
        if (r > 0) {
            k.x += 0x9E3779B9u;
            k.y += 0xBB67AE85u;
        }
        uint32_t hi0 = __umulhi(0xD2511F53u, c.x), lo0 = 0xD2511F53u * c.x;
        uint32_t hi1 = __umulhi(0xCD9E8D57u, c.z), lo1 = 0xCD9E8D57u * c.z;
        c = make_uint4(hi1 ^ c.y ^ k.x, lo1, hi0 ^ c.w ^ k.y, lo0);
    }
    return c;
}

enum : uint32_t { KIND_PHOTON = 1u, KIND_EYE = 2u, KIND_TRAIN = 3u };
// photon dims: 0 light pick, 1-2 position, 3-4 direction, 5-7 proposal,
// bounce b: 8+3b .. 10+3b.  eye dims: 0-1 jitter, bounce b: 2+3b .. 4+3b.
#define PH_DIM_BOUNCE 8
#define EYE_DIM_BOUNCE 2

struct Rng {
    uint2 key;
    uint4 ctr;
    uint4 buf;
    uint32_t block;

    __device__ inline void init(uint64_t seed, uint32_t kind, uint32_t pass, uint32_t index,
                                uint32_t scene) {
        key = make_uint2((uint32_t)seed, (uint32_t)(seed >> 32));
        ctr = make_uint4(index, scene, pass, kind << 24);
        block = 0xFFFFFFFFu;
    }
    // uniform of dimension `dim` in [0, 1): top 24 bits * 2^-24
    __device__ inline float at(int dim) {
        uint32_t blk = (uint32_t)dim >> 2;
        if (blk != block) {
            uint4 c = ctr;
            c.w |= blk;
            buf = philox4x32_10(c, key);
            block = blk;
        }
        uint32_t w = (dim & 3) == 0 ? buf.x : ((dim & 3) == 1 ? buf.y : ((dim & 3) == 2 ? buf.z : buf.w));
        return (float)(w >> 8) * (1.0f / 16777216.0f);
    }
};

// ---------------------------------------------------------------------------
// Octahedral unit-vector encoding, 2 x snorm16 packed in 32 bits.
// ---------------------------------------------------------------------------
__device__ inline float sgnf(float v) { return v < 0.0f ? -1.0f : 1.0f; }
__device__ inline uint32_t oct_encode(v3f n) {
    float s = fabsf(n.x) + fabsf(n.y) + fabsf(n.z);
    float x = n.x / s, y = n.y / s;
    if (n.z < 0.0f) {
        float ox = (1.0f - fabsf(y)) * sgnf(x);
        float oy = (1.0f - fabsf(x)) * sgnf(y);
        x = ox;
        y = oy;
    }
    int qx = (int)rintf(fminf(fmaxf(x, -1.0f), 1.0f) * 32767.0f);
    int qy = (int)rintf(fminf(fmaxf(y, -1.0f), 1.0f) * 32767.0f);
    return ((uint32_t)(qx & 0xFFFF)) | ((uint32_t)(qy & 0xFFFF) << 16);
}
__device__ inline v3f oct_decode(uint32_t e) {
    float x = (float)(int16_t)(e & 0xFFFF) * (1.0f / 32767.0f);
    float y = (float)(int16_t)(e >> 16) * (1.0f / 32767.0f);
    float z = 1.0f - fabsf(x) - fabsf(y);
    if (z < 0.0f) {
        float ox = (1.0f - fabsf(y)) * sgnf(x);
        float oy = (1.0f - fabsf(x)) * sgnf(y);
        x = ox;
        y = oy;
    }
    return normalize(mk(x, y, z));
}

// ---------------------------------------------------------------------------
// Device scene layout
// ---------------------------------------------------------------------------
// 4-wide BVH node, 128 bytes: child boxes in SoA FP32 (96 B) + 4 child refs.
// ref >= 0: inner node index.  ref < 0: leaf, ~ref = (first << 3) | (count - 1)
// over the leaf-ordered triangle array.  Empty slots have an inverted box.
struct alignas(128) Node4 {
    float lox[4], hix[4], loy[4], hiy[4], loz[4], hiz[4];
    int32_t child[4];
    int32_t pad[4];
};
static_assert(sizeof(Node4) == 128, "Node4 must be one 128-byte line");

// Triangles per BVH leaf at most (the leaf encoding holds up to 8).  On the
// caustic box, 1 and 2 trace equally fast and 4, 6, 8 slower (4.44, 4.45,
// 4.73, 4.97, 5.49 ms per photon pass); 2 keeps half the nodes of 1.
#define LEAF_MAX 2
#define EMPTY_REF 0x7FFFFFFF

__host__ __device__ inline int32_t leaf_ref(int32_t first, int32_t count) {
    return ~((first << 3) | (count - 1));
}

// Material in device form (48 bytes).
struct Mat {
    int32_t type;
    float rho[3];
    float le[3];
    float rr_mean;  // (rho.x + rho.y + rho.z) / 3: photon Russian roulette
    float rr_max;   // max(rho): eye Russian roulette
    float ior;
    int32_t sss;   // index into the dipole table (BSSRDF materials), else -1
    float pad;
};

// Dipole BSSRDF constants of one material (sources/bssrdf.cc:46-62), per
// channel: z+ = 1 / sigma'_t, z- = z+ (1 + 4A/3), sigma_tr = sqrt(3 sigma_a
// sigma'_t), k = alpha' / (4 pi sigma_tr); out = (1 - F_dr) / pi
struct DipoleDev {
    float zp[3], zn[3], tr[3], k[3];
    float out;
    float pad[3];
};

struct LightDev {
    v3f corner, eu, ev, normal;
    v3f flux;      // photon flux per emitted photon, set per render
    float cdf_hi;  // cumulative selection probability (last = 1)
};

struct SceneDev {
    const Node4* nodes;
    const float4* tris;      // [3 * n] leaf order: v0 (w = original id bits), v1, v2
    const int32_t* tri_mat;  // [n] original order
    const Mat* mats;
    const LightDev* lights;
    int32_t n_lights;
    int32_t root;            // root ref (inner index 0, a leaf ref, or EMPTY_REF)
    int32_t n_nodes;         // inner nodes (their bytes fit a 32-bit buffer offset)
    int64_t n_tris;
    // environment map (env null: black).  Lat-long texels, row 0 at +y; the
    // importance map's 64 x 64 equal-area cells with their probabilities and
    // running sums (FP64, as the host built them).
    const float4* env;
    const double* env_pdf;
    const double* env_cdf;
    int32_t env_w, env_h, env_last;
    float env_r;             // radius of the emission disc (the triangles' bounding sphere)
    v3f env_c;               // its centre
    double env_k;            // pi R^2 (4 pi / 4096) / (N P_env)
    double pick_total;       // emitted power of area lights + environment
    double pick_area;        // of the area lights: a photon whose light uniform u has u * total >= this
};

#define ENV_CELLS 64

// Envmap::sampleFromDir (sources/envmap.cc:47-60) in FP64: the texel of
// direction d, theta = acos(y) down the rows, phi = atan2(z, x) in [0, 2 pi)
// across the columns.
__device__ inline v3f env_lookup(const SceneDev& sc, double dx, double dy, double dz) {
    const double theta = acos(fmin(1.0, fmax(-1.0, dy)));
    double phi = atan2(dz, dx);
    if (phi < 0.0) phi += 2.0 * 3.14159265358979323846;
    int ix = (int)(phi / (2.0 * 3.14159265358979323846) * sc.env_w);
    int iy = (int)(theta / 3.14159265358979323846 * sc.env_h);
    ix = ix < 0 ? 0 : (ix > sc.env_w - 1 ? sc.env_w - 1 : ix);
    iy = iy < 0 ? 0 : (iy > sc.env_h - 1 ? sc.env_h - 1 : iy);
    const float4 t = sc.env[(int64_t)iy * sc.env_w + ix];
    return mk(t.x, t.y, t.z);
}

struct Hit {
    float t;
    int32_t tri;  // leaf-order triangle index, -1 = miss
    float u, v;
};

// Woop, Benthin, Wald (2013) watertight ray-triangle test in FP32.  No FP
// contraction here: the edge functions of two triangles sharing an edge must
// be exact negations of each other.
struct RayPre {
    v3f o;
    int kx, ky, kz;
    float sx, sy, sz;
};

__device__ inline RayPre ray_pre(v3f o, v3f d) {
    RayPre p;
    p.o = o;
    float ax = fabsf(d.x), ay = fabsf(d.y), az = fabsf(d.z);
    int kz = (ax > ay) ? (ax > az ? 0 : 2) : (ay > az ? 1 : 2);
    int kx = kz == 2 ? 0 : kz + 1;
    int ky = kx == 2 ? 0 : kx + 1;
    if (get(d, kz) < 0.0f) {
        int t = kx;
        kx = ky;
        ky = t;
    }
    p.kx = kx;
    p.ky = ky;
    p.kz = kz;
    float dz = get(d, kz);
    p.sx = get(d, kx) / dz;
    p.sy = get(d, ky) / dz;
    p.sz = 1.0f / dz;
    return p;
}

__device__ inline bool tri_watertight(const RayPre& p, float4 v0, float4 v1, float4 v2, float tmin,
                                      float tmax, float* t_out, float* u_out, float* v_out) {
#pragma clang fp contract(off)
    v3f A = xyz(v0) - p.o, B = xyz(v1) - p.o, C = xyz(v2) - p.o;
    float Akz = get(A, p.kz), Bkz = get(B, p.kz), Ckz = get(C, p.kz);
    float Ax = get(A, p.kx) - p.sx * Akz, Ay = get(A, p.ky) - p.sy * Akz;
    float Bx = get(B, p.kx) - p.sx * Bkz, By = get(B, p.ky) - p.sy * Bkz;
    float Cx = get(C, p.kx) - p.sx * Ckz, Cy = get(C, p.ky) - p.sy * Ckz;
    float U = Cx * By - Cy * Bx;
    float V = Ax * Cy - Ay * Cx;
    float W = Bx * Ay - By * Ax;
    if (U == 0.0f || V == 0.0f || W == 0.0f) {
        double CxBy = (double)Cx * (double)By, CyBx = (double)Cy * (double)Bx;
        U = (float)(CxBy - CyBx);
        double AxCy = (double)Ax * (double)Cy, AyCx = (double)Ay * (double)Cx;
        V = (float)(AxCy - AyCx);
        double BxAy = (double)Bx * (double)Ay, ByAx = (double)By * (double)Ax;
        W = (float)(BxAy - ByAx);
    }
    if ((U < 0.0f || V < 0.0f || W < 0.0f) && (U > 0.0f || V > 0.0f || W > 0.0f)) return false;
    float det = U + V + W;
    if (det == 0.0f) return false;
    float Az = p.sz * Akz, Bz = p.sz * Bkz, Cz = p.sz * Ckz;
    float T = U * Az + V * Bz + W * Cz;
    float sg = det < 0.0f ? -1.0f : 1.0f;
    float Ts = T * sg, dets = det * sg;
    if (Ts <= tmin * dets || Ts > tmax * dets) return false;
    float rcp = 1.0f / det;
    *t_out = T * rcp;
    *u_out = V * rcp;
    *v_out = W * rcp;
    return true;
}

// ---------------------------------------------------------------------------
// Wave-level helpers (wave64)
// ---------------------------------------------------------------------------
__device__ inline uint64_t lanemask_lt() {
    const int lane = threadIdx.x & 63;
    return lane == 0 ? 0ull : (~0ull >> (64 - lane));
}

__device__ inline unsigned long long wave_sum(unsigned long long v) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    return v;
}

// Kahan-compensated FP32 accumulation of one value into sum[k] / comp[k]
__device__ inline void kahan_add(float* sum, float* comp, int64_t k, float v) {
    float s = sum[k], c = comp[k];
    float y = v - c;
    float t = s + y;
    comp[k] = (t - s) - y;
    sum[k] = t;
}

// ---------------------------------------------------------------------------
// Reference hash grid cell and bucket (sources/hash_grid_detail.h:35-43, 60-62):
// |trunc((x - min) * scale)| in FP64, wrapping Teschner hash modulo the size
// ---------------------------------------------------------------------------
__device__ inline int32_t grid_cell(double x, double mn, double scale) {
#pragma clang fp contract(off)
    double b = (x - mn) * scale;
    int32_t c = (int32_t)b;
    return c < 0 ? -c : c;
}

__device__ inline uint32_t grid_hash(int32_t ix, int32_t iy, int32_t iz, uint32_t size) {
    uint32_t h = ((uint32_t)ix * 73856093u) ^ ((uint32_t)iy * 19349663u) ^ ((uint32_t)iz * 83492791u);
    return h % size;
}

// ---------------------------------------------------------------------------
// Learned emission proposal: the direction pdf about a light's normal w is the
// defensive mixture  p(w) = lambda cos(theta) / pi + (1 - lambda) q_k / dOmega
// over 32 x 32 equal-area cells (cos(theta) in 32 bands of 1/32, phi in 32
// sectors; dOmega = 2 pi / 1024 for every cell), with q the CDF row of the
// anchor (the light-position cell of the sample).  The cosine component is the
// plain emission of sources/sampler.cc:17-32, so lambda = 1 reproduces it bit
// for bit.  Photon dims: 3-4 direction, 5 component choice, 6-7 jitter inside
// the chosen cell.
// ---------------------------------------------------------------------------
#define PROP_SIDE 32
#define PROP_CELLS 1024
#define PH_DIM_PROP_SEL 5
#define PH_DIM_PROP_JIT 6

struct PropDev {
    const float* cdf;  // [side * side][1024] inclusive CDF per anchor, last entry 1
    float lambda;      // weight of the cosine component
    int32_t side;      // anchors per light edge; 0 = plain cosine emission
};

struct PropPick {
    uint32_t cell;     // direction cell kz * 32 + kphi
    float cos_t;       // cos(theta) of the direction
    float phi_u;       // phi / (2 pi); < 0: the cosine component (use on_hemisphere)
};

// Component, cell and (cos theta, phi) of an emitted direction; pu, pv are the
// light-sample coordinates along edge_u, edge_v.  r3, r4 are dims 3-4.
__device__ inline PropPick prop_pick(const PropDev& p, float pu, float pv, float r3, float r4,
                                     float sel, float j6, float j7) {
    PropPick k;
    if (p.side == 0 || sel < p.lambda) {
        // cosine component: phi = 2 pi r3, sin^2(theta) = r4
        k.cos_t = sqrtf(1.0f - r4);
        const int kz = min((int)((1.0f - k.cos_t) * (float)PROP_SIDE), PROP_SIDE - 1);
        const int kp = min((int)(r3 * (float)PROP_SIDE), PROP_SIDE - 1);
        k.cell = (uint32_t)(kz * PROP_SIDE + kp);
        k.phi_u = -1.0f;
        return k;
    }
    const int ia = min((int)(pu * (float)p.side), p.side - 1);
    const int ja = min((int)(pv * (float)p.side), p.side - 1);
    const float* cdf = p.cdf + (size_t)(ia * p.side + ja) * PROP_CELLS;
    int lo = 0, hi = PROP_CELLS - 1;  // first k with r3 < cdf[k]
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (r3 < cdf[mid]) hi = mid;
        else lo = mid + 1;
    }
    k.cell = (uint32_t)lo;
    const int kz = lo / PROP_SIDE, kp = lo % PROP_SIDE;
    k.cos_t = 1.0f - ((float)kz + j6) * (1.0f / PROP_SIDE);
    k.phi_u = ((float)kp + j7) * (1.0f / PROP_SIDE);
    return k;
}

// Flux factor (cos(theta) / pi) / p(w) of a picked direction; exactly 1 when
// lambda = 1 (or the proposal is off)
__device__ inline float prop_weight(const PropDev& p, float pu, float pv, const PropPick& k) {
    if (p.side == 0) return 1.0f;
    const int ia = min((int)(pu * (float)p.side), p.side - 1);
    const int ja = min((int)(pv * (float)p.side), p.side - 1);
    const float* cdf = p.cdf + (size_t)(ia * p.side + ja) * PROP_CELLS;
    const float q = cdf[k.cell] - (k.cell ? cdf[k.cell - 1] : 0.0f);
    return k.cos_t / (p.lambda * k.cos_t + (1.0f - p.lambda) * (0.5f * PROP_CELLS) * q);
}
