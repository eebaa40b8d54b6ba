// render.hip -- the probabilistic PPM pass on gfx950: photon pass, eye pass,
// fixed-radius gather with Kahan accumulation, and the small kernels between.
//
// Restates sources/ppm_probabilistic.cc:123-343 (tracePhotons, traceRays,
// radiance) in the P-KZ form: area-light emission, photons stored at
// Lambertian hits with Russian roulette by mean albedo, mirror / dielectric
// continuation, eye Russian roulette from bounce 3, one hit point per pixel
// per pass, and a fixed-radius cone-filtered density estimate.
#include "internal.h"
#include "traverse.h"

namespace {


// Append one record per active lane with one atomic per wave; returns the slot.
__device__ inline unsigned long long wave_append(unsigned long long* counter) {
    const uint64_t m = __ballot(1);
    const int leader = __ffsll((unsigned long long)m) - 1;
    const int lane = threadIdx.x & 63;
    unsigned long long base = 0;
    if (lane == leader) base = atomicAdd(counter, (unsigned long long)__popcll(m));
    base = __shfl(base, leader);
    return base + (unsigned long long)__popcll(m & lanemask_lt());
}


// sampler.cc:17-32 (cosine-weighted direction about w)
__device__ inline v3f on_hemisphere(v3f w, float r1, float r2) {
    v3f u = fabsf(w.x) > 1.0e-12f ? normalize(cross(mk(0.0f, 1.0f, 0.0f), w))
                                  : normalize(cross(mk(1.0f, 0.0f, 0.0f), w));
    v3f v = cross(w, u);
    float s, c;
    __sincosf(2.0f * PPPM_PI * r1, &s, &c);
    float z2s = sqrtf(r2);
    return normalize(u * (c * z2s) + v * (s * z2s) + w * sqrtf(1.0f - r2));
}

// Direction (cos theta, phi = 2 pi phi_u) in on_hemisphere's frame about w
__device__ inline v3f cell_direction(v3f w, float cos_t, float phi_u) {
    v3f u = fabsf(w.x) > 1.0e-12f ? normalize(cross(mk(0.0f, 1.0f, 0.0f), w))
                                  : normalize(cross(mk(1.0f, 0.0f, 0.0f), w));
    v3f v = cross(w, u);
    float s, c;
    __sincosf(2.0f * PPPM_PI * phi_u, &s, &c);
    const float sin_t = sqrtf(fmaxf(0.0f, 1.0f - cos_t * cos_t));
    return normalize(u * (c * sin_t) + v * (s * sin_t) + w * cos_t);
}

__device__ inline v3f reflect(v3f v, v3f n) { return v - n * (2.0f * dot(n, v)); }

// brdf.cc:117-142 with reflectance.h:7-34: returns the new direction, *pdf
__device__ inline v3f refraction_sample(v3f in, v3f n, float ior, float r1, float* pdf) {
    bool into = true;
    v3f on = n;
    if (dot(in, n) >= 0.0f) {
        into = false;
        on = -n;
    }
    v3f rd = reflect(in, n);
    float nnt = into ? 1.0f / ior : ior;
    float ddn = dot(in, on);
    float cos2t = 1.0f - nnt * nnt * (1.0f - ddn * ddn);
    if (cos2t < 0.0f) {
        *pdf = 1.0f;
        return rd;
    }
    v3f td = normalize(in * nnt - n * ((into ? 1.0f : -1.0f) * (ddn * nnt + sqrtf(cos2t))));
    float a = ior - 1.0f, b = ior + 1.0f;
    float R0 = (a * a) / (b * b);
    float c = 1.0f - (into ? -ddn : dot(td, -on));
    float c2 = c * c;
    float fre = R0 + (1.0f - R0) * (c2 * c2 * c);
    float ftr = 1.0f - fre;
    const float rp = 0.5f;  // 0.25 + 0.5 * REFLECT_PROBABILITY (brdf.cc:133)
    if (r1 < rp) {
        *pdf = rp / fre;
        return rd;
    }
    *pdf = (1.0f - rp) / ftr;
    return td;
}

__device__ inline v3f spawn(v3f p, v3f ng, v3f dir, float eps) {
    return dot(dir, ng) >= 0.0f ? p + ng * eps : p - ng * eps;
}

#define TRACE_TMAX 1.0e30f

// Unfinished rays a photon wave leaves in flight when it stops to shade and refill.

// Emission of photon i: light by power, uniform point, cosine direction (or
// the learned proposal's mixture); returns the ray origin, direction, flux.
// ENV: the scene may have an environment light (sc.env set).  Its FP64 code
// needs more registers than the rest of a kernel, so kernels of scenes
// without one are compiled without it; the numbers drawn are the same.
template <bool ENV>
__device__ inline void emit_photon(const PhotonPassArgs& a, const SceneDev& sc, Rng& rng, v3f& o, v3f& dir,
                                   v3f& flux) {
    float ul = rng.at(0);
    if (ENV && sc.env && (sc.n_lights == 0 || (double)ul * sc.pick_total >= sc.pick_area)) {
        // environment: a direction w by the importance cells (envmap.cc:62-80;
        // FP64, so the cell, the texel and the weight are the host's), a point
        // on a disc of the bounding sphere's radius facing w, travelling along
        // -w; flux = L(w) pi R^2 (4 pi / 4096) / (pdf(cell) N P_env)
        const double r = (double)rng.at(3);
        int lo = 0, hi = ENV_CELLS * ENV_CELLS;
        while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            if (sc.env_cdf[mid] < r) lo = mid + 1;
            else hi = mid;
        }
        const int idx = lo < ENV_CELLS * ENV_CELLS ? lo : sc.env_last;
        const double u = ((idx % ENV_CELLS) + (double)rng.at(4)) / ENV_CELLS;
        const double v = ((idx / ENV_CELLS) + (double)rng.at(5)) / ENV_CELLS;
        const double phi = u * 3.14159265358979323846 * 2.0;
        const double y = (1.0 - v) * 2.0 - 1.0;
        const double sy = sqrt(1.0 - y * y);
        const double wx = sy * cos(phi), wz = sy * sin(phi);
        const v3f L = env_lookup(sc, wx, y, wz);
        const float k = (float)(sc.env_k / sc.env_pdf[idx]);
        flux = L * k;
        const v3f w = mk((float)wx, (float)y, (float)wz);
        const v3f ta = fabsf(w.x) > 1.0e-12f ? normalize(cross(mk(0.0f, 1.0f, 0.0f), w))
                                            : normalize(cross(mk(1.0f, 0.0f, 0.0f), w));
        const v3f tb = cross(w, ta);
        const float rad = sc.env_r * sqrtf(rng.at(1));
        float sn, cs;
        __sincosf(2.0f * PPPM_PI * rng.at(2), &sn, &cs);
        const v3f pos = sc.env_c + w * (sc.env_r * 1.0001f) + ta * (rad * cs) + tb * (rad * sn);
        dir = -w;
        o = pos + dir * a.ray_eps;
        return;
    }
    int l = 0;
    while (l < sc.n_lights - 1 && sc.lights[l].cdf_hi <= ul) l++;
    const LightDev L = sc.lights[l];
    const float pu = rng.at(1), pv = rng.at(2);
    const v3f pos = L.corner + L.eu * pu + L.ev * pv;
    flux = L.flux;
    const float r3 = rng.at(3), r4 = rng.at(4);
    if (a.prop.side == 0) {
        dir = on_hemisphere(L.normal, r3, r4);
    } else {
        const float sel = rng.at(PH_DIM_PROP_SEL);
        const float j6 = rng.at(PH_DIM_PROP_JIT), j7 = rng.at(PH_DIM_PROP_JIT + 1);
        const PropPick k = prop_pick(a.prop, pu, pv, r3, r4, sel, j6, j7);
        dir = k.phi_u < 0.0f ? on_hemisphere(L.normal, r3, r4) : cell_direction(L.normal, k.cos_t, k.phi_u);
        flux = flux * prop_weight(a.prop, pu, pv, k);
    }
    o = pos + L.normal * a.ray_eps;
}

// One lane per photon path, with path regeneration at the level of the
// traversal: the wave traces until at most a.trace_keep (PHOTON_TRACE_KEEP by
// default) of its lanes still have a ray in flight, then the lanes whose ray is finished shade their
// hit and start their next segment, lanes whose path has ended take the next
// photon index from a pass-wide counter (one atomic per wave), and the
// traversal resumes with the unfinished rays where they stopped.  Lanes thus
// do not idle behind the longest ray of their wave.  Once the counter has run
// out every ray is traced to its end.  A photon's random numbers are keyed by
// its index, so its path and records do not depend on the lane that traces it.
// Four waves per SIMD: the 101 registers the kernel takes unspilled (without
// the environment-light code; with it, spills), measured faster in the
// pipelined passes than five waves at 96 registers.
template <bool COUNT, bool ENV>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) void k_photon(PhotonPassArgs a) {
    __shared__ int32_t s_stack[STACK_LDS * 256];
    int32_t* stk = s_stack + threadIdx.x;
    int32_t ovf_stack[STACK_OVF];
    unsigned long long segs = 0;
    bool ovf = false;
    const SceneDev& sc = a.sc;
    const uint32_t pass = a.pd ? a.pd->pass : a.pass;
    const int lane = threadIdx.x & 63;
    Rng rng;
    v3f o = mk(0.0f, 0.0f, 0.0f), dir = o, flux = o;
    uint32_t i = 0;
    int bounce = 0;
    uint32_t stored = 0;     // records the current photon has stored
    bool live = false;       // the lane holds a photon path
    bool exhausted = false;  // the pass-wide counter has run out (wave-uniform)
    TraceRay ray;
    TraceState ts;
    ts.cur = EMPTY_REF;
    ts.sp = 0;
    TravCount tc;
    while (true) {
        if (live && ts.cur == (int32_t)EMPTY_REF) {
            // the segment's ray is finished: shade its hit
            segs++;
            const Hit h = ts.h;
            if (h.tri < 0) {
                live = false;
            } else {
                const int dd = PH_DIM_BOUNCE + 3 * bounce;
                const float r0 = rng.at(dd), r1 = rng.at(dd + 1), r2 = rng.at(dd + 2);
                const Mat m = sc.mats[sc.tri_mat[tri_orig_id(sc, h.tri)]];
                const v3f hp = o + dir * h.t;
                const v3f ng = tri_normal(sc, h.tri);
                const v3f on = dot(ng, dir) < 0.0f ? ng : -ng;
                if (m.type & PPPM_BSDF_EMITTER) {
                    live = false;
                } else {
                    if ((m.type & PPPM_BSDF_BSSRDF) && a.irr_pos && dot(ng, dir) < 0.0f) {
                        // irradiance photon (arriving from outside) for the
                        // subsurface estimate; the path goes on through the
                        // dielectric boundary below
                        unsigned long long slot = wave_append(&a.counters->irr_stored);
                        if (slot < (unsigned long long)a.irr_cap) {
                            a.irr_pos[slot] = make_float4(hp.x, hp.y, hp.z, __uint_as_float(oct_encode(on)));
                            a.irr_flux[slot] = make_float4(flux.x, flux.y, flux.z,
                                                           __uint_as_float((i << 6) | (uint32_t)bounce));
                        }
                    }
                    if (m.type & PPPM_BSDF_LAMBERTIAN) {
                        unsigned long long slot = wave_append(&a.counters->photons_stored);
                        if (slot < (unsigned long long)a.cap) {
                            a.ph_pos[slot * a.rec_stride] = make_float4(hp.x, hp.y, hp.z, __uint_as_float(oct_encode(on)));
                            a.ph_flux[slot * a.rec_stride] = make_float4(flux.x, flux.y, flux.z,
                                                          __uint_as_float((i << 6) | (uint32_t)bounce));
                            if (a.rec_rank) a.rec_rank[slot] = (uint8_t)stored;
                        }
                        stored++;
                        if (r0 < m.rr_mean) {
                            dir = on_hemisphere(on, r1, r2);
                            o = spawn(hp, ng, dir, a.ray_eps);
                            flux = flux * mk(m.rho[0], m.rho[1], m.rho[2]) * (1.0f / m.rr_mean);
                        } else {
                            live = false;
                        }
                    } else {
                        float pdf = 1.0f;
                        v3f nd;
                        if (m.type & PPPM_BSDF_SPECULAR) nd = reflect(dir, on);
                        else nd = refraction_sample(dir, ng, m.ior, r0, &pdf);
                        o = spawn(hp, ng, nd, a.ray_eps);
                        dir = nd;
                        flux = flux * mk(m.rho[0], m.rho[1], m.rho[2]) * (1.0f / pdf);
                    }
                    bounce++;
                }
            }
            if (live && (bounce >= a.max_bounces || vmax3(flux) <= 0.0f)) live = false;
            if (live) trace_start(sc, o, dir, 0.0f, TRACE_TMAX, ray, ts);
            else if (a.rec_count) a.rec_count[i] = stored;  // the path has ended
        }
        if (!exhausted) {
            const uint64_t need = __ballot(!live);
            if (need != 0) {
                const int leader = __ffsll((unsigned long long)need) - 1;
                unsigned long long base = 0;
                if (lane == leader) base = atomicAdd(&a.counters->photon_next, (unsigned long long)__popcll(need));
                base = __shfl(base, leader);
                if (!live) {
                    const unsigned long long k = base + (unsigned long long)__popcll(need & lanemask_lt());
                    if (k < (unsigned long long)a.n_photons) {
                        i = (uint32_t)k;
                        rng.init(a.seed, KIND_PHOTON, pass, i, a.scene_id);
                        emit_photon<ENV>(a, sc, rng, o, dir, flux);
                        bounce = 0;
                        stored = 0;
                        live = a.max_bounces > 0 && vmax3(flux) > 0.0f;
                        if (live) trace_start(sc, o, dir, 0.0f, TRACE_TMAX, ray, ts);
                        else if (a.rec_count) a.rec_count[i] = 0u;
                    }
                }
                exhausted = base + (unsigned long long)__popcll(need) >= (unsigned long long)a.n_photons;
            }
        }
        if (__ballot(live) == 0) {
            if (exhausted) break;
            continue;
        }
        trace_resume<256, COUNT>(sc, ray, ts, stk, ovf_stack, &ovf, &tc, exhausted ? 0 : a.trace_keep);
    }
    segs = wave_sum(segs);
    uint64_t any_ovf = __ballot(ovf);
    if (lane == 0) {
        atomicAdd(&a.counters->photon_segments, segs);
        if (any_ovf) atomicAdd(&a.counters->stack_overflow, 1ull);
    }
    if (COUNT) {
        const unsigned long long nn = wave_sum((unsigned long long)tc.nodes);
        const unsigned long long nt = wave_sum((unsigned long long)tc.tris);
        if (lane == 0) {
            atomicAdd(&a.counters->photon_nodes, nn);
            atomicAdd(&a.counters->photon_tris, nt);
        }
    }
}


// One lane per pixel path (a grid-stride loop past the grid limit).  Not
// persistent on purpose: the eye pass runs beside the gather of an earlier
// pass, and its blocks, which end at different times, hand their registers
// back to the gather as they end; a persistent eye kernel with path
// regeneration was faster alone (2.8 against 3.2 ms) but held its registers
// for the whole pass and slowed the pipelined passes by 1 %.
template <bool COUNT, bool ENV>
__global__ __launch_bounds__(256) void k_eye(EyePassArgs a) {
    __shared__ int32_t s_stack[STACK_LDS * 256];
    int32_t* stk = s_stack + threadIdx.x;
    unsigned long long segs = 0, nhit = 0;
    bool ovf = false;
    TravCount tc;
    const SceneDev& sc = a.sc;
    const int W = a.cam.width, H = a.cam.height;
    const int64_t np = (int64_t)W * H;
    const v3f C = mk(a.cam_basis[0], a.cam_basis[1], a.cam_basis[2]);
    const v3f D = mk(a.cam_basis[3], a.cam_basis[4], a.cam_basis[5]);
    const v3f U = mk(a.cam_basis[6], a.cam_basis[7], a.cam_basis[8]);
    const v3f V = mk(a.cam_basis[9], a.cam_basis[10], a.cam_basis[11]);
    for (int64_t p = (int64_t)blockIdx.x * 256 + threadIdx.x; p < np;
         p += (int64_t)gridDim.x * 256) {
        const int x = (int)(p % W), row = (int)(p / W);
        Rng rng;
        rng.init(a.seed, KIND_EYE, a.pd ? a.pd->pass : a.pass, (uint32_t)p, a.scene_id);
        const float px = (float)x + rng.at(0);
        const float py = (float)(H - 1 - row) + rng.at(1);
        v3f d = normalize(D + U * (2.0f * px / (float)W - 1.0f) + V * (2.0f * py / (float)H - 1.0f));
        v3f o = C;
        v3f thr = mk(1.0f, 1.0f, 1.0f);
        v3f direct = mk(0.0f, 0.0f, 0.0f);
        float4 hit_pos = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        float4 hit_w = make_float4(0.0f, 0.0f, 0.0f, __uint_as_float(NO_HITPOINT));
        // leaving the scene, or at the bounce limit: the environment seen along
        // d, or black (ppm_probabilistic.cc:248-251), looked up once after the
        // loop; kernels built with ENV false (scenes without an environment)
        // leave the lookup out: its FP64 angles take the kernel from 105 to
        // 177 registers
        bool escaped = false;
        for (int bounce = 0;; bounce++) {
            if (bounce >= a.max_bounces) {
                escaped = true;
                break;
            }
            Hit h = trace_closest<256, COUNT>(sc, o, d, 0.0f, TRACE_TMAX, stk, &ovf, &tc);
            segs++;
            if (h.tri < 0) {
                escaped = true;
                break;
            }
            const int dd = EYE_DIM_BOUNCE + 3 * bounce;
            const float r0 = rng.at(dd), r1 = rng.at(dd + 1), r2 = rng.at(dd + 2);
            (void)r2;
            const Mat m = sc.mats[sc.tri_mat[tri_orig_id(sc, h.tri)]];
            const v3f hp = o + d * h.t;
            const v3f ng = tri_normal(sc, h.tri);
            if (m.type & PPPM_BSDF_EMITTER) {
                if (dot(d, ng) < 0.0f) direct = thr * mk(m.le[0], m.le[1], m.le[2]);
                break;
            }
            float roulette = m.rr_max;
            if (a.roulette && bounce >= 3) {
                if (r0 > roulette) break;
            } else {
                roulette = 1.0f;
            }
            const v3f on = dot(ng, d) < 0.0f ? ng : -ng;
            const v3f rho = mk(m.rho[0], m.rho[1], m.rho[2]);
            if (m.type & PPPM_BSDF_LAMBERTIAN) {
                const v3f w = thr * rho * (PPPM_INV_PI / roulette);
                hit_pos = make_float4(hp.x, hp.y, hp.z, __uint_as_float(oct_encode(on)));
                hit_w = make_float4(w.x, w.y, w.z, __uint_as_float((uint32_t)p));
                nhit++;
                break;
            }
            float pdf = 1.0f;
            v3f nd;
            if (m.type & PPPM_BSDF_SPECULAR) nd = reflect(d, on);
            else nd = refraction_sample(d, ng, m.ior, r1, &pdf);
            if ((m.type & PPPM_BSDF_BSSRDF) && a.sss_pos && (dot(nd, ng) < 0.0f) == (dot(d, ng) < 0.0f)) {
                // transmitted into the translucent body: the path ends in the
                // dipole estimate at hp, weighted F_t / (1 - P) (ppm_probabilistic.cc:272-296)
                const v3f w = thr * (1.0f / (pdf * roulette));
                unsigned long long slot = wave_append(&a.counters->sss_hits);
                if (slot < (unsigned long long)a.sss_cap) {
                    a.sss_pos[slot] = make_float4(hp.x, hp.y, hp.z, __int_as_float(m.sss));
                    a.sss_w[slot] = make_float4(w.x, w.y, w.z, __uint_as_float((uint32_t)p));
                }
                break;
            }
            thr = thr * rho * (1.0f / (pdf * roulette));
            o = spawn(hp, ng, nd, a.ray_eps);
            d = nd;
        }
        if (ENV && escaped && sc.env) direct = thr * env_lookup(sc, d.x, d.y, d.z);
        // the hit point of pixel p lives in slot p (pixel id all ones: none), so
        // the hit point order entering the Morton sort is the pixel order; a
        // slot without hit point carries the emission the path saw directly
        // (a path ends at a hit point or at an emitter, never both)
        if (__float_as_uint(hit_w.w) == NO_HITPOINT) hit_w = make_float4(direct.x, direct.y, direct.z, hit_w.w);
        a.hp_pos[p] = hit_pos;
        a.hp_w[p] = hit_w;
        if (a.img_sum && (direct.x != 0.0f || direct.y != 0.0f || direct.z != 0.0f)) {
            kahan_add(a.img_sum, a.img_comp, 3 * p + 0, direct.x);
            kahan_add(a.img_sum, a.img_comp, 3 * p + 1, direct.y);
            kahan_add(a.img_sum, a.img_comp, 3 * p + 2, direct.z);
        }
    }
    segs = wave_sum(segs);
    nhit = wave_sum(nhit);
    uint64_t any_ovf = __ballot(ovf);
    if ((threadIdx.x & 63) == 0) {
        atomicAdd(&a.counters->eye_segments, segs);
        atomicAdd(&a.counters->hitpoints, nhit);
        if (any_ovf) atomicAdd(&a.counters->stack_overflow, 1ull);
    }
    if (COUNT) {
        const unsigned long long nn = wave_sum((unsigned long long)tc.nodes);
        const unsigned long long nt = wave_sum((unsigned long long)tc.tris);
        if ((threadIdx.x & 63) == 0) {
            atomicAdd(&a.counters->eye_nodes, nn);
            atomicAdd(&a.counters->eye_tris, nt);
        }
    }
}

// The direct emission of an eye pass run with img_sum = null, added from the
// pixel slots later and in pass order: the same Kahan additions, in the same
// order per pixel, as the eye kernel makes when it adds them itself.
__global__ void k_direct_add(const float4* __restrict__ hp_w, int64_t np, float* img_sum, float* img_comp) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= np) return;
    const float4 w = hp_w[p];
    if (__float_as_uint(w.w) != NO_HITPOINT) return;
    if (w.x != 0.0f || w.y != 0.0f || w.z != 0.0f) {
        kahan_add(img_sum, img_comp, 3 * p + 0, w.x);
        kahan_add(img_sum, img_comp, 3 * p + 1, w.y);
        kahan_add(img_sum, img_comp, 3 * p + 2, w.z);
    }
}

// The image average of pppm_image_read ((sum - comp) + (dsum - dcomp)) / passes
// in FP64, encoded per pixel: fmt 0 = 8-bit RGB after c^ginv (ginv <= 0: none)
// clamped and truncated; fmt 1 = RGBE (frexp of the largest channel);
// fmt 2 = FP32 RGB.
__global__ void k_image_encode(const float* __restrict__ s, const float* __restrict__ c,
                               const float* __restrict__ ds, const float* __restrict__ dc, int64_t npix, double inv,
                               int fmt, double ginv, uint8_t* out8, float* outf) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= npix) return;
    double v[3];
    for (int k = 0; k < 3; k++) {
        const int64_t j = 3 * i + k;
        v[k] = (((double)s[j] - (double)c[j]) + ((double)ds[j] - (double)dc[j])) * inv;
    }
    if (fmt == 0) {
        for (int k = 0; k < 3; k++) {
            double d = ginv > 0.0 ? pow(v[k], ginv) : v[k];
            d = d < 0.0 ? 0.0 : (d > 1.0 ? 1.0 : d);
            out8[3 * i + k] = (uint8_t)(255.0 * d);
        }
    } else if (fmt == 1) {
        for (int k = 0; k < 3; k++) v[k] = v[k] > 0.0 ? v[k] : 0.0;
        double d = fmax(v[0], fmax(v[1], v[2]));
        uint8_t* o = out8 + 4 * i;
        if (!(d > 1.0e-32)) {
            o[0] = o[1] = o[2] = o[3] = 0;
        } else {
            int e;
            const double m = frexp(d, &e);
            d = m * 256.0 / d;
            o[0] = (uint8_t)(v[0] * d);
            o[1] = (uint8_t)(v[1] * d);
            o[2] = (uint8_t)(v[2] * d);
            o[3] = (uint8_t)(e + 128);
        }
    } else {
        for (int k = 0; k < 3; k++) outf[3 * i + k] = (float)v[k];
    }
}

// Accumulator probe: lane k Kahan-adds v[(i + k) % n] for i = 0 .. n-1 into
// sum[k] / comp[k] with the image accumulator's own kahan_add.
__global__ void k_kahan_probe(const float* __restrict__ v, int64_t n, int lanes, float* sum, float* comp) {
    const int k = threadIdx.x;
    if (k >= lanes) return;
    sum[k] = 0.0f;
    comp[k] = 0.0f;
    for (int64_t i = 0; i < n; i++) kahan_add(sum, comp, k, v[(i + k) % n]);
}

// Sort key of a photon record: its photon index.  A photon's records are
// appended by the lane that traces it, in bounce order, at increasing slots,
// so the stable sort by index alone orders them by (index, bounce): the
// bounce's six bits need no radix pass.
__global__ void k_origin_keys(const float4* __restrict__ ph_flux, int64_t n, uint32_t* keys,
                              uint32_t* vals, const unsigned long long* count, uint32_t sentinel) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const bool real = !count || i < (int64_t)min((unsigned long long)n, *count);
    keys[i] = real ? __float_as_uint(ph_flux[i].w) >> 6 : sentinel;
    vals[i] = (uint32_t)i;
}

__global__ void k_pass_advance(PassDev* pd, const double* radii) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    const uint32_t p = pd->pass + 1;
    const double r = radii[p - 1];
    pd->pass = p;
    pd->radius = r;
    pd->scale = 1.0 / (2.0 * r);
}

__global__ void k_capacity_check(PassDev* pd, const Counters* c, int64_t cap) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    if ((int64_t)c->photons_stored > cap || (int64_t)c->irr_stored > cap) pd->overflow = 1u;
}

__global__ void k_permute2(const float4* __restrict__ a_in, const float4* __restrict__ b_in,
                           const uint32_t* __restrict__ perm, const uint32_t* __restrict__ perm2,
                           int64_t n, float4* __restrict__ a_out, float4* __restrict__ b_out, int stride) {
    int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    uint32_t j = perm2 ? perm2[k] : (uint32_t)k;
    uint32_t src = perm ? perm[j] : j;
    a_out[k] = a_in[(int64_t)src * stride];
    b_out[k] = b_in[(int64_t)src * stride];
}

__global__ void k_compose(const uint32_t* perm1, const uint32_t* perm2, int64_t n, uint32_t* out) {
    int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) out[k] = perm1[perm2[k]];
}

__global__ void k_zero(uint32_t* p, int64_t words) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < words;
         i += (int64_t)gridDim.x * blockDim.x)
        p[i] = 0u;
}

// [n][3] float arrays -> float4 records (w = oct normal bits when normals != 0)
__global__ void k_pack(const float* pos3, const float* b3, const float* n3, int64_t n, float4* pos4,
                       float4* b4, int normals) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint32_t e = 0;
    if (normals) e = oct_encode(normalize(mk(n3[3 * i], n3[3 * i + 1], n3[3 * i + 2])));
    pos4[i] = make_float4(pos3[3 * i], pos3[3 * i + 1], pos3[3 * i + 2], __uint_as_float(e));
    b4[i] = make_float4(b3[3 * i], b3[3 * i + 1], b3[3 * i + 2], __uint_as_float((uint32_t)i));
}

}  // namespace

static inline unsigned grid1(int64_t n, int b) { return (unsigned)((n + b - 1) / b); }

void launch_photon_pass(hipStream_t s, const PhotonPassArgs& a, int grid_limit) {
    if (a.n_photons <= 0) return;
    unsigned g = grid1(a.n_photons, 256);
    if (grid_limit > 0 && g > (unsigned)grid_limit) g = (unsigned)grid_limit;
    const bool env = a.sc.env != nullptr;
    if (a.count_traversal && env)
        hipLaunchKernelGGL((k_photon<true, true>), dim3(g), dim3(256), 0, s, a);
    else if (a.count_traversal)
        hipLaunchKernelGGL((k_photon<true, false>), dim3(g), dim3(256), 0, s, a);
    else if (env)
        hipLaunchKernelGGL((k_photon<false, true>), dim3(g), dim3(256), 0, s, a);
    else
        hipLaunchKernelGGL((k_photon<false, false>), dim3(g), dim3(256), 0, s, a);
}

void launch_eye_pass(hipStream_t s, const EyePassArgs& a, int grid_limit) {
    int64_t np = (int64_t)a.cam.width * a.cam.height;
    if (np <= 0) return;
    unsigned g = grid1(np, 256);
    if (grid_limit > 0 && g > (unsigned)grid_limit) g = (unsigned)grid_limit;
    const bool env = a.sc.env != nullptr;
    if (a.count_traversal && env)
        hipLaunchKernelGGL((k_eye<true, true>), dim3(g), dim3(256), 0, s, a);
    else if (a.count_traversal)
        hipLaunchKernelGGL((k_eye<true, false>), dim3(g), dim3(256), 0, s, a);
    else if (env)
        hipLaunchKernelGGL((k_eye<false, true>), dim3(g), dim3(256), 0, s, a);
    else
        hipLaunchKernelGGL((k_eye<false, false>), dim3(g), dim3(256), 0, s, a);
}

void launch_direct_add(hipStream_t s, const float4* hp_w, int64_t np, float* img_sum, float* img_comp) {
    if (np <= 0) return;
    hipLaunchKernelGGL(k_direct_add, dim3(grid1(np, 256)), dim3(256), 0, s, hp_w, np, img_sum, img_comp);
}

void launch_image_encode(hipStream_t s, const float* sum, const float* comp, const float* dsum, const float* dcomp,
                         int64_t npix, double inv, int fmt, double ginv, uint8_t* out8, float* outf) {
    if (npix <= 0) return;
    hipLaunchKernelGGL(k_image_encode, dim3(grid1(npix, 256)), dim3(256), 0, s, sum, comp, dsum, dcomp, npix, inv, fmt,
                       ginv, out8, outf);
}

void launch_kahan_probe(hipStream_t s, const float* v, int64_t n, int lanes, float* sum, float* comp) {
    hipLaunchKernelGGL(k_kahan_probe, dim3(1), dim3(64), 0, s, v, n, lanes, sum, comp);
}

void launch_origin_keys(hipStream_t s, const float4* ph_flux, int64_t n, uint32_t* keys,
                        uint32_t* vals, const unsigned long long* count, uint32_t sentinel) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_origin_keys, dim3(grid1(n, 256)), dim3(256), 0, s, ph_flux, n, keys, vals, count, sentinel);
}

void launch_pass_advance(hipStream_t s, PassDev* pd, const double* radii) {
    hipLaunchKernelGGL(k_pass_advance, dim3(1), dim3(64), 0, s, pd, radii);
}

void launch_capacity_check(hipStream_t s, PassDev* pd, const Counters* counters, int64_t cap) {
    hipLaunchKernelGGL(k_capacity_check, dim3(1), dim3(64), 0, s, pd, counters, cap);
}

void launch_permute2(hipStream_t s, const float4* a_in, const float4* b_in, const uint32_t* perm,
                     const uint32_t* perm2, int64_t n, float4* a_out, float4* b_out, int stride) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_permute2, dim3(grid1(n, 256)), dim3(256), 0, s, a_in, b_in, perm, perm2, n,
                       a_out, b_out, stride);
}

void launch_compose(hipStream_t s, const uint32_t* perm1, const uint32_t* perm2, int64_t n,
                    uint32_t* out) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_compose, dim3(grid1(n, 256)), dim3(256), 0, s, perm1, perm2, n, out);
}

void launch_zero_u32(hipStream_t s, void* p, int64_t words) {
    if (words <= 0) return;
    int64_t g = grid1(words, 256);
    if (g > 4096) g = 4096;
    hipLaunchKernelGGL(k_zero, dim3((unsigned)g), dim3(256), 0, s, (uint32_t*)p, words);
}

void launch_pack_points(hipStream_t s, const float* pos3, const float* a3, const float* n3,
                        int64_t n, float4* pos4, float4* b4, int normals) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_pack, dim3(grid1(n, 256)), dim3(256), 0, s, pos3, a3, n3, n, pos4, b4,
                       normals);
}

namespace {
__global__ __launch_bounds__(256) void k_stream_copy(const float4* __restrict__ src,
                                                     float4* __restrict__ dst, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256)
        dst[i] = src[i];
}
}  // namespace

void launch_stream_copy(hipStream_t s, const float4* src, float4* dst, int64_t n) {
    if (n <= 0) return;
    int64_t g = grid1(n, 256);
    if (g > 65536) g = 65536;
    hipLaunchKernelGGL(k_stream_copy, dim3((unsigned)g), dim3(256), 0, s, src, dst, n);
}
