/*
 * oracle.c -- CPU restatement of the tatsy-pppm probabilistic PPM pass (P-KZ).
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP path and the CPU
 * baseline timed by bench.py.  Never linked into the product.
 *
 * Function                      restates                                   pinned by
 * ----------------------------  -----------------------------------------  -----------------------------------------
 * orc_philox4x32_10             Philox-4x32-10 (Salmon et al. 2011)         Random123 known-answer vectors
 * tri_intersect                 sources/triangle.cc:53-78                  closed forms; compiled reference (_ref)
 * qbvh_build / qbvh_intersect   sources/qbvh_accel.cc:139-285              brute force (tests/test_trimesh.cc:62-117
 *                                                                          strategy); compiled reference (_ref)
 * orc_hash / orc_hashgrid_*     sources/hash_grid_detail.h:26-72           compiled reference HashGrid<int> (_ref);
 *                                                                          brute-force bucket lists
 * orc_radius_schedule           sources/ppm_probabilistic.cc:94 (literal)  Gamma-function closed form
 *                               + Knaus-Zwicker 2011 shrink
 * on_hemisphere                 sources/sampler.cc:17-32                   compiled reference (_ref); cosine moments
 * total_reflection              sources/reflectance.h:7-34                 compiled reference (_ref); Fresnel 0.04
 * bsdf_sample                   sources/brdf.cc:23-26, 49-52, 117-142      compiled reference (_ref)
 * camera_ray                    sources/perspective_camera.cc:42-46        centre/corner closed forms
 * orc_trace_photons             sources/ppm_probabilistic.cc:123-208       first-deposit flux = pi Le A / N exactly;
 *                                                                          mean stores 1 / (1 - rho) in a closed box
 * orc_eye_pass (eye_path)       sources/ppm_probabilistic.cc:210-343       mirror-corridor closed form: hit point by
 *                                                                          unfolding, weight prod(rho) / pi / roulettes,
 *                                                                          emitter seen through mirrors
 * orc_gather                    sources/ppm_probabilistic.cc:300-334       brute force; uniform-density closed form
 * orc_render / orc_render_ex    sources/ppm_probabilistic.cc:24-121        sphere furnace L = rho Phi / (pi A (1 - rho));
 *                                                                          the independent path tracer on the box scene
 * orc_render_ex (subsurface)    ppm_probabilistic.cc:272-298;              translucent plane under a parallel light:
 *                               subsurface_integrator.cc:234-238, 329-365  rho F_r Le + F_t (1 - F_dr) Le I_Rd / pi
 * orc_pathtrace                 none (next-event path tracer; the          furnace closed form; agreement with the
 *                               reference's path_tracing.cc:136-198        photon mapping render
 *                               cannot see area lights)
 * orc_dipole / orc_dipole_sum   sources/bssrdf.cc:33-104;                  compiled reference DipoleBSSRDF (_ref golden,
 *                               subsurface_integrator.cc:147-167, 234-238  fast exp); closed forms of R_d at d = 0 and
 *                                                                          of its clamped area integral
 * emit_direction /              no reference counterpart (the learned      uniform-table and one-hot-table closed
 *   orc_trace_photons_prop        mixture replacing sampler.cc:17-32)        forms; lambda = 1 equals on_hemisphere;
 *                                                                          E[factor] = 1 (unbiasedness)
 * orc_dipole_octree             sources/subsurface_integrator.cc:69-167    compiled reference Octree (_ref golden, three
 *                                                                          point sets); max_error 0 = the exact sum
 * orc_scene_set_envmap /        sources/envmap.cc:47-163                   compiled reference Envmap (_ref): importance
 *   orc_envmap_lookup /                                                    cells and running sums bit for bit, lookups,
 *   orc_envmap_sample_dir                                                  photon directions; env-lit floor closed form
 * orc_render_hachisuka          sources/progressive_photon_mapping.cc:     sphere furnace closed form; agreement with
 *                               22-287                                     orc_render and orc_pathtrace on the box
 *
 * Deliberate differences from the literal reference (each is a quirk the
 * reference cannot run with): fixed-radius gather instead of the swapped kNN,
 * Knaus-Zwicker radius shrink, area-light emission, photon/hit-point normal
 * agreement test, counter-based RNG, ray origins offset by ray_eps along the
 * surface normal.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* sources/common.h:12-18 */
#define INFTY 1.0e20
#define EPS 1.0e-12
#define IOR_OBJECT 1.5
#define IOR_VACCUM 1.0
#define REFLECT_PROBABILITY 0.5
static const double PI = 3.14159265358979323846;

/* ------------------------------------------------------------------------ */
/* Random numbers                                                            */
/* ------------------------------------------------------------------------ */

void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int r = 0; r < 10; r++) {
        if (r > 0) { k0 += 0x9E3779B9u; k1 += 0xBB67AE85u; }
        uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        uint32_t n0 = hi1 ^ c1 ^ k0;
        uint32_t n2 = hi0 ^ c3 ^ k1;
        c0 = n0; c1 = lo1; c2 = n2; c3 = lo0;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

/* Stream of uniforms: key = seed, counter = (index, scene, pass, kind<<24 | block).
 * Uniform = top 24 bits * 2^-24, exact in FP32 and FP64. */
typedef struct {
    uint32_t key[2];
    uint32_t ctr[4];
    uint32_t buf[4];
    int pos;
    uint32_t block;
} rng_t;

static void rng_init(rng_t* r, uint64_t seed, uint32_t kind, uint32_t pass, uint32_t index,
                     uint32_t scene) {
    r->key[0] = (uint32_t)seed;
    r->key[1] = (uint32_t)(seed >> 32);
    r->ctr[0] = index;
    r->ctr[1] = scene;
    r->ctr[2] = pass;
    r->ctr[3] = kind << 24;
    r->pos = 4;
    r->block = 0;
}

static double rng_at(rng_t* r, int dim) {
    uint32_t blk = (uint32_t)dim >> 2;
    if (r->pos == 4 || blk != r->block) {
        uint32_t c[4] = {r->ctr[0], r->ctr[1], r->ctr[2], r->ctr[3] | blk};
        orc_philox4x32_10(c, r->key, r->buf);
        r->block = blk;
        r->pos = 0;
    }
    return (double)(r->buf[dim & 3] >> 8) * (1.0 / 16777216.0);
}

void orc_uniforms(uint64_t seed, uint32_t kind, uint32_t pass, uint32_t index, uint32_t scene,
                  int32_t n, double* out) {
    rng_t r;
    rng_init(&r, seed, kind, pass, index, scene);
    for (int i = 0; i < n; i++) out[i] = rng_at(&r, i);
}

#define KIND_PHOTON 1u
#define KIND_EYE 2u
/* photon dims: 0 light pick, 1-2 position, 3-4 direction, 5-7 reserved,
 * bounce b uses 8+3b .. 10+3b.  eye dims: 0-1 jitter, bounce b 2+3b .. 4+3b. */
#define PH_DIM_BOUNCE 8
#define EYE_DIM_BOUNCE 2

/* ------------------------------------------------------------------------ */
/* Vector3D (sources/vector3d.cc)                                            */
/* ------------------------------------------------------------------------ */

typedef struct { double x, y, z; } v3;
static inline v3 V(double x, double y, double z) { v3 r = {x, y, z}; return r; }
static inline v3 add(v3 a, v3 b) { return V(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline v3 sub(v3 a, v3 b) { return V(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline v3 mul(v3 a, v3 b) { return V(a.x * b.x, a.y * b.y, a.z * b.z); }
static inline v3 scl(v3 a, double s) { return V(a.x * s, a.y * s, a.z * s); }
static inline v3 neg(v3 a) { return V(-a.x, -a.y, -a.z); }
/* vector3d.cc:44-53 */
static inline double dot(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline v3 cross(v3 u, v3 v) {
    return V(u.y * v.z - u.z * v.y, u.z * v.x - u.x * v.z, u.x * v.y - u.y * v.x);
}
static inline double norm(v3 a) { return sqrt(dot(a, a)); }
/* vector3d.cc:63-65: divides by norm + EPS (multiplying by 1/(n+EPS), :144-150) */
static inline v3 normalized(v3 a) { return scl(a, 1.0 / (norm(a) + EPS)); }
/* vector3d.cc:67-69 */
static inline v3 reflect(v3 v, v3 n) { return sub(v, scl(n, 2.0 * dot(n, v))); }
static inline double vmax3(v3 a) { return fmax(a.x, fmax(a.y, a.z)); }
static inline double vget(v3 a, int d) { return d == 0 ? a.x : (d == 1 ? a.y : a.z); }

/* ------------------------------------------------------------------------ */
/* Scene + QBVH                                                              */
/* ------------------------------------------------------------------------ */

typedef struct {
    float lo[3][4], hi[3][4];   /* childBoxes[min|max][x|y|z][4] (qbvh_accel.h:17) */
    int32_t child[4];           /* node index, or -1 */
    int32_t tri_start, tri_count; /* leaf range in tri_order */
    int8_t sep[3];
    int8_t is_leaf;
} qnode;

struct orc_scene {
    int64_t n;
    v3* p0; v3* p1; v3* p2;
    int32_t* mat;
    orc_bsdf* bsdf;
    int32_t nb;
    orc_light* lights;
    int32_t nl;
    double* light_cdf;
    double light_total;
    qnode* nodes;
    int64_t n_nodes, cap_nodes;
    int32_t* tri_order;
    int32_t depth;
    v3 bmin, bmax;
    /* environment map (env_w == 0: none) */
    int32_t env_w, env_h, env_last;
    double* env_rgb;        /* [h][w][3] */
    double env_pdf[4096], env_cdf[4096];
    double env_power;       /* selection weight against the area lights */
    v3 env_c;               /* bounding sphere of the triangles */
    double env_r;
};

/* triangle.cc:53-78 (Moller-Trumbore, FP64).  Returns t or -1. */
static double tri_intersect(const orc_scene* s, int64_t i, v3 o, v3 d) {
    v3 e1 = sub(s->p1[i], s->p0[i]);
    v3 e2 = sub(s->p2[i], s->p0[i]);
    v3 pv = cross(d, e2);
    double det = dot(e1, pv);
    if (det > -EPS && det < EPS) return -1.0;
    double invdet = 1.0 / det;
    v3 tv = sub(o, s->p0[i]);
    double u = dot(tv, pv) * invdet;
    if (u < 0.0 || u > 1.0) return -1.0;
    v3 qv = cross(tv, e1);
    double v = dot(d, qv) * invdet;
    if (v < 0.0 || u + v > 1.0) return -1.0;
    double t = dot(e2, qv) * invdet;
    return t > EPS ? t : -1.0;
}

/* triangle.cc:45-47 */
static v3 tri_normal(const orc_scene* s, int64_t i) {
    return normalized(cross(sub(s->p1[i], s->p0[i]), sub(s->p2[i], s->p0[i])));
}

static double* g_sort_key; /* centroid coordinate for the comparator */
static int cmp_centroid(const void* a, const void* b) {
    int32_t ia = *(const int32_t*)a, ib = *(const int32_t*)b;
    double ka = g_sort_key[ia], kb = g_sort_key[ib];
    if (ka < kb) return -1;
    if (ka > kb) return 1;
    return ia < ib ? -1 : (ia > ib ? 1 : 0); /* deterministic ties (std::sort's are unspecified) */
}

static double* g_centroid[3];

static int64_t qnode_alloc(orc_scene* s) {
    if (s->n_nodes == s->cap_nodes) {
        s->cap_nodes = s->cap_nodes ? s->cap_nodes * 2 : 1024;
        s->nodes = (qnode*)realloc(s->nodes, sizeof(qnode) * (size_t)s->cap_nodes);
    }
    memset(&s->nodes[s->n_nodes], 0, sizeof(qnode));
    return s->n_nodes++;
}

static void range_box(const orc_scene* s, const int32_t* idx, int64_t n, v3* lo, v3* hi) {
    *lo = V(INFTY, INFTY, INFTY);
    *hi = V(-INFTY, -INFTY, -INFTY);
    for (int64_t k = 0; k < n; k++) {
        int32_t i = idx[k];
        v3 ps[3] = {s->p0[i], s->p1[i], s->p2[i]};
        for (int j = 0; j < 3; j++) {
            lo->x = fmin(lo->x, ps[j].x); lo->y = fmin(lo->y, ps[j].y); lo->z = fmin(lo->z, ps[j].z);
            hi->x = fmax(hi->x, ps[j].x); hi->y = fmax(hi->y, ps[j].y); hi->z = fmax(hi->z, ps[j].z);
        }
    }
}

/* qbvh_accel.cc:150-203: 4-way object-median split, <= 3 triangles per leaf. */
static int64_t qbvh_build(orc_scene* s, int32_t* idx, int64_t n, int dim, int depth) {
    int64_t me = qnode_alloc(s);
    if (depth > s->depth) s->depth = depth;
    if (n <= 3) {
        s->nodes[me].is_leaf = 1;
        s->nodes[me].tri_start = (int32_t)(idx - s->tri_order);
        s->nodes[me].tri_count = (int32_t)n;
        for (int c = 0; c < 4; c++) s->nodes[me].child[c] = -1;
        return me;
    }
    g_sort_key = g_centroid[dim];
    qsort(idx, (size_t)n, sizeof(int32_t), cmp_centroid);
    int64_t mid = n / 2;
    g_sort_key = g_centroid[(dim + 1) % 3];
    qsort(idx, (size_t)mid, sizeof(int32_t), cmp_centroid);
    qsort(idx + mid, (size_t)(n - mid), sizeof(int32_t), cmp_centroid);
    int64_t b[5] = {0, mid / 2, mid, mid + mid / 2, n};
    for (int c = 0; c < 4; c++) {
        v3 lo, hi;
        range_box(s, idx + b[c], b[c + 1] - b[c], &lo, &hi);
        /* boxes stored as float (qbvh_accel.cc:182-183) */
        s->nodes[me].lo[0][c] = (float)lo.x; s->nodes[me].lo[1][c] = (float)lo.y; s->nodes[me].lo[2][c] = (float)lo.z;
        s->nodes[me].hi[0][c] = (float)hi.x; s->nodes[me].hi[1][c] = (float)hi.y; s->nodes[me].hi[2][c] = (float)hi.z;
    }
    for (int c = 0; c < 4; c++) {
        int64_t ch = qbvh_build(s, idx + b[c], b[c + 1] - b[c], (dim + 2) % 3, depth + 1);
        s->nodes[me].child[c] = (int32_t)ch;
    }
    s->nodes[me].sep[0] = (int8_t)dim;
    s->nodes[me].sep[1] = (int8_t)((dim + 1) % 3);
    s->nodes[me].sep[2] = (int8_t)((dim + 1) % 3);
    s->nodes[me].is_leaf = 0;
    return me;
}

/* qbvh_accel.cc:205-285.  Near-first child order by entry distance replaces
 * the sign-bit order table (it only changes which of two equal-t hits wins). */
static int32_t qbvh_intersect(const orc_scene* s, v3 o, v3 d, double tmin_ray, double tmax_ray,
                              double* t_out) {
    float org[3] = {(float)o.x, (float)o.y, (float)o.z};
    float idir[3];
    for (int k = 0; k < 3; k++) {
        double dk = vget(d, k);
        idir[k] = dk == 0.0 ? (float)1.0e20 : (float)(1.0 / dk);
    }
    int sgn[3] = {idir[0] > 0.0f ? 0 : 1, idir[1] > 0.0f ? 0 : 1, idir[2] > 0.0f ? 0 : 1};
    double best = tmax_ray;
    int32_t hit = -1;
    int32_t stack[256];
    int sp = 0;
    if (s->n_nodes == 0) { *t_out = best; return -1; }
    stack[sp++] = 0;
    while (sp > 0) {
        const qnode* nd = &s->nodes[stack[--sp]];
        if (nd->is_leaf) {
            for (int k = 0; k < nd->tri_count; k++) {
                int32_t ti = s->tri_order[nd->tri_start + k];
                double t = tri_intersect(s, ti, o, d);
                if (t > tmin_ray && best > t) { best = t; hit = ti; }
            }
            continue;
        }
        float tn[4], tf[4];
        for (int c = 0; c < 4; c++) { tn[c] = 0.0f; tf[c] = (float)best; }
        for (int k = 0; k < 3; k++) {
            for (int c = 0; c < 4; c++) {
                float a = ((sgn[k] ? nd->hi[k][c] : nd->lo[k][c]) - org[k]) * idir[k];
                float b = ((sgn[k] ? nd->lo[k][c] : nd->hi[k][c]) - org[k]) * idir[k];
                tn[c] = fmaxf(tn[c], a);
                tf[c] = fminf(tf[c], b);
            }
        }
        int order[4], m = 0;
        for (int c = 0; c < 4; c++) if (tf[c] >= tn[c]) order[m++] = c;
        /* push far-to-near so that the nearest is popped first */
        for (int a = 1; a < m; a++)
            for (int b = a; b > 0 && tn[order[b]] > tn[order[b - 1]]; b--) {
                int tmp = order[b]; order[b] = order[b - 1]; order[b - 1] = tmp;
            }
        for (int c = 0; c < m; c++) stack[sp++] = nd->child[order[c]];
    }
    *t_out = best;
    return hit;
}

static int32_t brute_intersect(const orc_scene* s, v3 o, v3 d, double tmin_ray, double tmax_ray,
                               double* t_out) {
    double best = tmax_ray;
    int32_t hit = -1;
    for (int64_t i = 0; i < s->n; i++) {
        double t = tri_intersect(s, i, o, d);
        if (t > tmin_ray && best > t) { best = t; hit = (int32_t)i; }
    }
    *t_out = best;
    return hit;
}

orc_scene* orc_scene_create(const float* tri, const int32_t* bsdf_id, int64_t n,
                            const orc_bsdf* table, int32_t nb, const orc_light* lights, int32_t nl) {
    orc_scene* s = (orc_scene*)calloc(1, sizeof(orc_scene));
    s->n = n;
    s->p0 = (v3*)malloc(sizeof(v3) * (size_t)(n ? n : 1));
    s->p1 = (v3*)malloc(sizeof(v3) * (size_t)(n ? n : 1));
    s->p2 = (v3*)malloc(sizeof(v3) * (size_t)(n ? n : 1));
    s->mat = (int32_t*)malloc(sizeof(int32_t) * (size_t)(n ? n : 1));
    s->bmin = V(INFTY, INFTY, INFTY);
    s->bmax = V(-INFTY, -INFTY, -INFTY);
    for (int64_t i = 0; i < n; i++) {
        const float* t = tri + 9 * i;
        s->p0[i] = V(t[0], t[1], t[2]);
        s->p1[i] = V(t[3], t[4], t[5]);
        s->p2[i] = V(t[6], t[7], t[8]);
        s->mat[i] = bsdf_id ? bsdf_id[i] : 0;
        for (int j = 0; j < 9; j++) {
            double c = t[j];
            int d = j % 3;
            if (d == 0) { s->bmin.x = fmin(s->bmin.x, c); s->bmax.x = fmax(s->bmax.x, c); }
            if (d == 1) { s->bmin.y = fmin(s->bmin.y, c); s->bmax.y = fmax(s->bmax.y, c); }
            if (d == 2) { s->bmin.z = fmin(s->bmin.z, c); s->bmax.z = fmax(s->bmax.z, c); }
        }
    }
    s->nb = nb;
    s->bsdf = (orc_bsdf*)malloc(sizeof(orc_bsdf) * (size_t)(nb ? nb : 1));
    if (nb) memcpy(s->bsdf, table, sizeof(orc_bsdf) * (size_t)nb);
    s->nl = nl;
    s->lights = (orc_light*)malloc(sizeof(orc_light) * (size_t)(nl ? nl : 1));
    s->light_cdf = (double*)malloc(sizeof(double) * (size_t)(nl + 1));
    if (nl) memcpy(s->lights, lights, sizeof(orc_light) * (size_t)nl);
    s->light_cdf[0] = 0.0;
    for (int l = 0; l < nl; l++) {
        const orc_light* L = &s->lights[l];
        v3 eu = V(L->edge_u[0], L->edge_u[1], L->edge_u[2]);
        v3 ev = V(L->edge_v[0], L->edge_v[1], L->edge_v[2]);
        double area = norm(cross(eu, ev));
        double lum = (L->emission[0] + (double)L->emission[1] + L->emission[2]) / 3.0;
        s->light_cdf[l + 1] = s->light_cdf[l] + PI * area * lum;
    }
    s->light_total = s->light_cdf[nl];
    /* QBVH build */
    s->tri_order = (int32_t*)malloc(sizeof(int32_t) * (size_t)(n ? n : 1));
    for (int d = 0; d < 3; d++) g_centroid[d] = (double*)malloc(sizeof(double) * (size_t)(n ? n : 1));
    for (int64_t i = 0; i < n; i++) {
        s->tri_order[i] = (int32_t)i;
        /* Triangle::gravity, triangle.cc:49-51 */
        v3 g = scl(add(add(s->p0[i], s->p1[i]), s->p2[i]), 1.0 / 3.0);
        g_centroid[0][i] = g.x; g_centroid[1][i] = g.y; g_centroid[2][i] = g.z;
    }
    if (n > 0) qbvh_build(s, s->tri_order, n, 0, 1);
    for (int d = 0; d < 3; d++) { free(g_centroid[d]); g_centroid[d] = NULL; }
    return s;
}

void orc_scene_destroy(orc_scene* s) {
    if (!s) return;
    free(s->env_rgb);
    free(s->p0); free(s->p1); free(s->p2); free(s->mat); free(s->bsdf);
    free(s->lights); free(s->light_cdf); free(s->nodes); free(s->tri_order);
    free(s);
}

int32_t orc_qbvh_stats(const orc_scene* s, int64_t* nodes, int64_t* leaves, int32_t* depth) {
    int64_t nl = 0;
    for (int64_t i = 0; i < s->n_nodes; i++) nl += s->nodes[i].is_leaf;
    if (nodes) *nodes = s->n_nodes;
    if (leaves) *leaves = nl;
    if (depth) *depth = s->depth;
    return 0;
}

static int32_t scene_intersect(const orc_scene* s, v3 o, v3 d, double* t) {
    return qbvh_intersect(s, o, d, 0.0, INFTY, t);
}

/* ------------------------------------------------------------------------ */
/* Environment map (sources/envmap.cc)                                       */
/* ------------------------------------------------------------------------ */

#define ENV_CELLS 64  /* IMPORTANCE_MAP_SIZE, envmap.h:31 */

/* Envmap::sampleFromDir (envmap.cc:47-60): nearest texel of the lat-long
 * image, theta = acos(y) from +y down the rows, phi = atan2(z, x) in [0, 2 pi)
 * across the columns */
static v3 env_lookup(const orc_scene* s, v3 d) {
    double theta = acos(d.y);
    double phi = atan2(d.z, d.x);
    if (phi < 0.0) phi += 2.0 * PI;
    double x = phi / (2.0 * PI) * s->env_w;
    double y = theta / PI * s->env_h;
    int ix = (int)x, iy = (int)y;
    ix = ix < 0 ? 0 : (ix > s->env_w - 1 ? s->env_w - 1 : ix);
    iy = iy < 0 ? 0 : (iy > s->env_h - 1 ? s->env_h - 1 : iy);
    const double* p = s->env_rgb + 3 * ((size_t)iy * s->env_w + ix);
    return V(p[0], p[1], p[2]);
}

/* the direction of importance-map coordinates (u, v) (envmap.cc:76-80):
 * equal-area cells, y = 1 - 2 v, phi = 2 pi u */
static v3 env_direction(double u, double v) {
    double phi = u * PI * 2.0;
    double y = (1.0 - v) * 2.0 - 1.0;
    return V(sqrt(1.0 - y * y) * cos(phi), y, sqrt(1.0 - y * y) * sin(phi));
}

int32_t orc_scene_set_envmap(orc_scene* s, const float* rgb, int32_t w, int32_t h) {
    free(s->env_rgb);
    s->env_rgb = NULL;
    s->env_w = s->env_h = 0;
    if (!rgb || w < ENV_CELLS || h < ENV_CELLS) return -1;
    s->env_rgb = (double*)malloc(sizeof(double) * 3 * (size_t)w * h);
    for (size_t i = 0; i < 3 * (size_t)w * h; i++) s->env_rgb[i] = rgb[i];
    s->env_w = w;
    s->env_h = h;
    /* Envmap::createImportanceMap, the warped part (envmap.cc:128-162): per
     * 64 x 64 cell the squared luminance of the mean of (W/64) x (H/64)
     * lookups, normalised by (total + EPS), and its running sum */
    const int sx_n = w / ENV_CELLS, sy_n = h / ENV_CELLS;
    double total = 0.0;
    for (int iy = 0; iy < ENV_CELLS; iy++)
        for (int ix = 0; ix < ENV_CELLS; ix++) {
            v3 acc = V(0, 0, 0);
            for (int sy = 0; sy < sy_n; sy++)
                for (int sx = 0; sx < sx_n; sx++) {
                    double u = (ix + (double)sx / sx_n) / ENV_CELLS;
                    double v = (iy + (double)sy / sy_n) / ENV_CELLS;
                    acc = add(acc, scl(env_lookup(s, env_direction(u, v)), 1.0 / (sx_n * sy_n)));
                }
            double l = acc.x * 0.2126 + acc.y * 0.7152 + acc.z * 0.0722;
            s->env_pdf[iy * ENV_CELLS + ix] = l * l;
            total += l * l;
        }
    s->env_pdf[0] /= (total + EPS);
    s->env_cdf[0] = s->env_pdf[0];
    for (int i = 1; i < ENV_CELLS * ENV_CELLS; i++) {
        s->env_pdf[i] /= (total + EPS);
        s->env_cdf[i] = s->env_pdf[i] + s->env_cdf[i - 1];
    }
    s->env_last = 0;
    for (int i = 0; i < ENV_CELLS * ENV_CELLS; i++)
        if (s->env_pdf[i] > 0.0) s->env_last = i;
    /* photons leave a disc as wide as the triangles' bounding sphere */
    s->env_c = scl(add(s->bmin, s->bmax), 0.5);
    s->env_r = 0.5 * norm(sub(s->bmax, s->bmin));
    /* selection weight: pi R^2 times the integral of the mean channel over
     * the sphere, texel by texel (texel solid angle (2 pi / w)(cos t0 - cos t1)) */
    double pw = 0.0;
    for (int iy = 0; iy < h; iy++) {
        double dom = (2.0 * PI / w) * (cos(PI * iy / h) - cos(PI * (iy + 1) / h));
        double row = 0.0;
        for (int ix = 0; ix < w; ix++) {
            const double* p = s->env_rgb + 3 * ((size_t)iy * w + ix);
            row += (p[0] + p[1] + p[2]) / 3.0;
        }
        pw += row * dom;
    }
    s->env_power = PI * s->env_r * s->env_r * pw;
    return 0;
}

/* first cell whose running sum reaches r (std::lower_bound, envmap.cc:67-69);
 * above the last sum (at most EPS below 1) the last cell with mass */
static int env_cell(const orc_scene* s, double r) {
    int lo = 0, hi = ENV_CELLS * ENV_CELLS;
    while (lo < hi) {
        int mid = (lo + hi) >> 1;
        if (s->env_cdf[mid] < r) lo = mid + 1;
        else hi = mid;
    }
    return lo < ENV_CELLS * ENV_CELLS ? lo : s->env_last;
}

void orc_envmap_cells(const orc_scene* s, double* pdf, double* cdf) {
    memcpy(pdf, s->env_pdf, sizeof(s->env_pdf));
    memcpy(cdf, s->env_cdf, sizeof(s->env_cdf));
}

void orc_envmap_lookup(const orc_scene* s, const double* dirs, int64_t n, double* rgb) {
    for (int64_t k = 0; k < n; k++) {
        v3 c = env_lookup(s, V(dirs[3 * k], dirs[3 * k + 1], dirs[3 * k + 2]));
        rgb[3 * k] = c.x;
        rgb[3 * k + 1] = c.y;
        rgb[3 * k + 2] = c.z;
    }
}

void orc_envmap_sample_dir(const orc_scene* s, const double* r, int64_t n, double* dir, double* l_over_pdf,
                           int32_t* cell) {
    for (int64_t k = 0; k < n; k++) {
        int idx = env_cell(s, r[3 * k]);
        double u = ((idx % ENV_CELLS) + r[3 * k + 1]) / ENV_CELLS;
        double v = ((idx / ENV_CELLS) + r[3 * k + 2]) / ENV_CELLS;
        v3 w = env_direction(u, v);
        v3 L = env_lookup(s, w);
        dir[3 * k] = w.x;
        dir[3 * k + 1] = w.y;
        dir[3 * k + 2] = w.z;
        l_over_pdf[3 * k] = L.x / s->env_pdf[idx];
        l_over_pdf[3 * k + 1] = L.y / s->env_pdf[idx];
        l_over_pdf[3 * k + 2] = L.z / s->env_pdf[idx];
        if (cell) cell[k] = idx;
    }
}

void orc_intersect_batch(const orc_scene* s, const float* rays, int64_t n, int32_t brute,
                         float* t_out, int32_t* tri_out) {
#pragma omp parallel for schedule(dynamic, 256)
    for (int64_t i = 0; i < n; i++) {
        const float* r = rays + 8 * i;
        v3 o = V(r[0], r[1], r[2]);
        v3 d = V(r[4], r[5], r[6]);
        double t;
        int32_t h = brute ? brute_intersect(s, o, d, r[3], r[7], &t)
                          : qbvh_intersect(s, o, d, r[3], r[7], &t);
        tri_out[i] = h;
        t_out[i] = h >= 0 ? (float)t : r[7];
    }
}

/* ------------------------------------------------------------------------ */
/* HashGrid (hash_grid_detail.h)                                             */
/* ------------------------------------------------------------------------ */

/* hash_grid_detail.h:60-63.  The int32 products wrap (two's complement), as
 * they do on every platform the reference targets; done here in uint32. */
uint32_t orc_hash(int32_t ix, int32_t iy, int32_t iz, int32_t hash_size) {
    uint32_t h = ((uint32_t)ix * 73856093u) ^ ((uint32_t)iy * 19349663u) ^ ((uint32_t)iz * 83492791u);
    return h % (uint32_t)hash_size;
}

/* hash_grid_detail.h:35-43: |trunc((x - min) * scale)| in FP64 */
static inline int32_t cell_coord(double x, double mn, double scale) {
    double b = (x - mn) * scale;
    int32_t c = (int32_t)b;
    return c < 0 ? -c : c;
}

static void buckets_to_ranges(const uint32_t* h, int64_t n, int32_t S, int32_t* sorted,
                              int32_t* cell_start, int32_t* cell_end) {
    /* stable counting sort by bucket: bucket h holds items in insertion order */
    int64_t* cnt = (int64_t*)calloc((size_t)S + 1, sizeof(int64_t));
    for (int64_t i = 0; i < n; i++) cnt[h[i] + 1]++;
    for (int32_t b = 0; b < S; b++) cnt[b + 1] += cnt[b];
    for (int32_t b = 0; b < S; b++) { cell_start[b] = (int32_t)cnt[b]; cell_end[b] = (int32_t)cnt[b + 1]; }
    int64_t* pos = (int64_t*)malloc(sizeof(int64_t) * ((size_t)S + 1));
    memcpy(pos, cnt, sizeof(int64_t) * ((size_t)S + 1));
    for (int64_t i = 0; i < n; i++) sorted[pos[h[i]]++] = (int32_t)i;
    free(pos);
    free(cnt);
}

int32_t orc_hashgrid_points(const float* pos, int64_t n, int32_t S, double scale,
                            const double bmin[3], int32_t* sorted, int32_t* cell_start,
                            int32_t* cell_end) {
    if (S <= 0) return -1;
    uint32_t* h = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)(n ? n : 1));
    for (int64_t i = 0; i < n; i++) {
        int32_t ix = cell_coord((double)pos[3 * i + 0], bmin[0], scale);
        int32_t iy = cell_coord((double)pos[3 * i + 1], bmin[1], scale);
        int32_t iz = cell_coord((double)pos[3 * i + 2], bmin[2], scale);
        h[i] = orc_hash(ix, iy, iz, S);
    }
    buckets_to_ranges(h, n, S, sorted, cell_start, cell_end);
    free(h);
    return 0;
}

int64_t orc_hashgrid_boxes(const double* box_min, const double* box_max, int64_t n, int32_t S,
                           double scale, const double bmin[3], int32_t* sorted_items, int64_t cap,
                           int32_t* cell_start, int32_t* cell_end) {
    if (S <= 0) return -1;
    /* hash_grid_detail.h:44-50: z outer, y, x inner, inclusive bounds */
    int64_t total = 0;
    for (int64_t i = 0; i < n; i++) {
        int32_t lo[3], hi[3];
        for (int d = 0; d < 3; d++) {
            lo[d] = cell_coord(box_min[3 * i + d], bmin[d], scale);
            hi[d] = cell_coord(box_max[3 * i + d], bmin[d], scale);
        }
        if (hi[0] >= lo[0] && hi[1] >= lo[1] && hi[2] >= lo[2])
            total += (int64_t)(hi[0] - lo[0] + 1) * (hi[1] - lo[1] + 1) * (hi[2] - lo[2] + 1);
    }
    if (total > cap) return -1;
    uint32_t* h = (uint32_t*)calloc((size_t)(total ? total : 1), sizeof(uint32_t));
    int32_t* item = (int32_t*)malloc(sizeof(int32_t) * (size_t)(total ? total : 1));
    int64_t k = 0;
    for (int64_t i = 0; i < n; i++) {
        int32_t lo[3], hi[3];
        for (int d = 0; d < 3; d++) {
            lo[d] = cell_coord(box_min[3 * i + d], bmin[d], scale);
            hi[d] = cell_coord(box_max[3 * i + d], bmin[d], scale);
        }
        for (int32_t iz = lo[2]; iz <= hi[2]; iz++)
            for (int32_t iy = lo[1]; iy <= hi[1]; iy++)
                for (int32_t ix = lo[0]; ix <= hi[0]; ix++) {
                    h[k] = orc_hash(ix, iy, iz, S);
                    item[k] = (int32_t)i;
                    k++;
                }
    }
    int32_t* perm = (int32_t*)malloc(sizeof(int32_t) * (size_t)(total ? total : 1));
    buckets_to_ranges(h, total, S, perm, cell_start, cell_end);
    for (int64_t j = 0; j < total; j++) sorted_items[j] = item[perm[j]];
    free(perm); free(item); free(h);
    return total;
}

/* ------------------------------------------------------------------------ */
/* Radius schedule                                                           */
/* ------------------------------------------------------------------------ */

void orc_radius_schedule(double r1, double alpha, int32_t n, int32_t literal, double* out) {
    if (n <= 0) return;
    out[0] = r1;
    if (literal) {
        /* ppm_probabilistic.cc:94: _radius = (t + 1.0) / (t + ALPHA) * _radius */
        double r = r1;
        for (int t = 1; t < n; t++) { r = (t + 1.0) / (t + alpha) * r; out[t] = r; }
    } else {
        /* Knaus & Zwicker 2011: r2_{i+1} = r2_i * ((i + alpha) / (i + 1)) */
        double r2 = r1 * r1;
        for (int i = 1; i < n; i++) { r2 = r2 * ((i + alpha) / (i + 1.0)); out[i] = sqrt(r2); }
    }
}

/* ------------------------------------------------------------------------ */
/* Sampling and BSDFs                                                        */
/* ------------------------------------------------------------------------ */

/* sampler.cc:17-32 */
static v3 on_hemisphere(v3 w, double r1, double r2) {
    v3 u;
    if (fabs(w.x) > EPS) u = normalized(cross(V(0.0, 1.0, 0.0), w));
    else u = normalized(cross(V(1.0, 0.0, 0.0), w));
    v3 v = cross(w, u);
    double t = 2.0 * PI * r1;
    double z2s = sqrt(r2);
    return normalized(add(add(scl(u, cos(t) * z2s), scl(v, sin(t) * z2s)), scl(w, sqrt(1.0 - r2))));
}

/* reflectance.h:7-34 */
static int total_reflection(int into, v3 in, v3 n, v3 on, double ior, v3* refl, v3* refr, double* fre,
                            double* ftr) {
    *refl = reflect(in, n);
    double nnt = into ? IOR_VACCUM / ior : ior / IOR_VACCUM;
    double ddn = dot(in, on);
    double cos2t = 1.0 - nnt * nnt * (1.0 - ddn * ddn);
    if (cos2t < 0.0) {
        *refr = V(0, 0, 0);
        *fre = 1.0;
        *ftr = 0.0;
        return 1;
    }
    *refr = normalized(sub(scl(in, nnt), scl(n, (into ? 1.0 : -1.0) * (ddn * nnt + sqrt(cos2t)))));
    double a = ior - IOR_VACCUM, b = ior + IOR_VACCUM;
    double R0 = (a * a) / (b * b);
    double c = 1.0 - (into ? -ddn : dot(*refr, neg(on)));
    *fre = R0 + (1.0 - R0) * (c * c * c * c * c);
    *ftr = 1.0 - *fre;
    return 0;
}

/* brdf.cc:23-26 (Lambert), 49-52 (mirror), 117-142 (refraction).  pdf keeps
 * the caller's 1.0 unless the BSDF sets it. */
static v3 bsdf_sample(const orc_bsdf* b, v3 in, v3 n, double r1, double r2, double* pdf) {
    if (b->type & ORC_LAMBERTIAN) {
        v3 on = dot(in, n) < 0.0 ? n : neg(n);
        return on_hemisphere(on, r1, r2);
    }
    if (b->type & ORC_SPECULAR) {
        v3 on = dot(in, n) < 0.0 ? n : neg(n);
        return reflect(in, on);
    }
    /* ORC_REFRACTION */
    int into = 1;
    v3 on = n;
    if (dot(in, n) >= 0.0) { into = 0; on = neg(n); }
    v3 rd, td;
    double fre, ftr;
    /* the material's index (IOR_OBJECT, common.h:16, when it names none) */
    double ior = b->ior > 0.0f ? (double)b->ior : IOR_OBJECT;
    if (total_reflection(into, in, n, on, ior, &rd, &td, &fre, &ftr)) {
        *pdf = 1.0;
        return rd;
    }
    double rp = 0.25 + 0.5 * REFLECT_PROBABILITY;
    if (r1 < rp) { *pdf = rp / fre; return rd; }
    *pdf = (1.0 - rp) / ftr;
    return td;
}

static inline v3 refl_of(const orc_bsdf* b) {
    return V(b->reflectance[0], b->reflectance[1], b->reflectance[2]);
}

/* Origin of a continuation ray: the hit point moved by eps to the side of the
 * surface the new direction leaves into. */
static inline v3 spawn(v3 p, v3 ng, v3 dir, double eps) {
    return dot(dir, ng) >= 0.0 ? add(p, scl(ng, eps)) : sub(p, scl(ng, eps));
}

/* ------------------------------------------------------------------------ */
/* Photon pass (ppm_probabilistic.cc:123-208, area-light emission)           */
/* ------------------------------------------------------------------------ */

typedef struct { float p[3], f[3], n[3]; int32_t bounce; } deposit;

/* Direction cell of the learned proposal (pppm_device.h prop_pick/prop_weight):
 * 32 x 32 equal-area cells about the light normal, 1 - cos(theta) in bands of
 * 1/32, phi in sectors of 2 pi / 32; mixture lambda * cosine + (1 - lambda) *
 * the anchor's CDF row; returns the flux factor (cos / pi) / p(w). */
static double emit_direction(const orc_proposal* pr, v3 nl, double pu, double pv, rng_t* r, v3* dir,
                             uint32_t* cell_out) {
    double r3 = rng_at(r, 3), r4 = rng_at(r, 4);
    if (!pr || pr->side == 0) {
        *dir = on_hemisphere(nl, r3, r4);
        return 1.0;
    }
    double sel = rng_at(r, 5), j6 = rng_at(r, 6), j7 = rng_at(r, 7);
    int side = pr->side;
    int ia = (int)(pu * side), ja = (int)(pv * side);
    if (ia > side - 1) ia = side - 1;
    if (ja > side - 1) ja = side - 1;
    const float* cdf = pr->cdf + (size_t)(ia * side + ja) * 1024;
    double cos_t;
    int kz, kp;
    if (sel < pr->lambda) {
        *dir = on_hemisphere(nl, r3, r4);
        cos_t = sqrt(1.0 - r4);
        kz = (int)((1.0 - cos_t) * 32.0);
        kp = (int)(r3 * 32.0);
        if (kz > 31) kz = 31;
        if (kp > 31) kp = 31;
    } else {
        int lo = 0, hi = 1023;
        while (lo < hi) {
            int mid = (lo + hi) >> 1;
            if (r3 < (double)cdf[mid]) hi = mid;
            else lo = mid + 1;
        }
        kz = lo / 32;
        kp = lo % 32;
        cos_t = 1.0 - ((double)kz + j6) / 32.0;
        double phi = 2.0 * PI * ((double)kp + j7) / 32.0;
        double sin_t = sqrt(fmax(0.0, 1.0 - cos_t * cos_t));
        v3 u;
        if (fabs(nl.x) > EPS) u = normalized(cross(V(0.0, 1.0, 0.0), nl));
        else u = normalized(cross(V(1.0, 0.0, 0.0), nl));
        v3 v = cross(nl, u);
        *dir = normalized(add(add(scl(u, cos(phi) * sin_t), scl(v, sin(phi) * sin_t)), scl(nl, cos_t)));
    }
    int cell = kz * 32 + kp;
    if (cell_out) *cell_out = (uint32_t)cell;
    float qf = cell ? cdf[cell] - cdf[cell - 1] : cdf[cell];
    return cos_t / (pr->lambda * cos_t + (1.0 - pr->lambda) * 512.0 * (double)qf);
}

/* Emission: pick a light by power, uniform point on the parallelogram,
 * cosine direction about its normal (sampler.cc:17-32), or the learned
 * mixture when pr is given; flux = P_total / N times the mixture factor. */
static void emit_photon(const orc_scene* s, const orc_proposal* pr, rng_t* r, int64_t N, v3* pos, v3* dir,
                        v3* flux, v3* nl_out) {
    double ul = rng_at(r, 0);
    double env_p = s->env_w > 0 ? s->env_power : 0.0;
    double total = s->light_total + env_p;
    double target = ul * total;
    if (env_p > 0.0 && (s->nl == 0 || target >= s->light_cdf[s->nl])) {
        /* environment: a direction w by the importance map (envmap.cc:62-80),
         * then a point on a disc of the bounding sphere's radius facing w,
         * travelling along -w; flux = L(w) pi R^2 / (p(w) N P_env) with the
         * solid-angle density p(w) = pdf(cell) 4096 / (4 pi) of the equal-area
         * cells (the reference's flux scale, :81-83, has no disc area) */
        int idx = env_cell(s, rng_at(r, 3));
        double u = ((idx % ENV_CELLS) + rng_at(r, 4)) / ENV_CELLS;
        double v = ((idx / ENV_CELLS) + rng_at(r, 5)) / ENV_CELLS;
        v3 w = env_direction(u, v);
        v3 L = env_lookup(s, w);
        double k = PI * s->env_r * s->env_r * (4.0 * PI / (ENV_CELLS * ENV_CELLS)) /
                   (s->env_pdf[idx] * (double)N * (env_p / total));
        *flux = scl(L, k);
        v3 a;
        if (fabs(w.x) > EPS) a = normalized(cross(V(0.0, 1.0, 0.0), w));
        else a = normalized(cross(V(1.0, 0.0, 0.0), w));
        v3 b = cross(w, a);
        double rad = s->env_r * sqrt(rng_at(r, 1)), ang = 2.0 * PI * rng_at(r, 2);
        *pos = add(add(s->env_c, scl(w, s->env_r * 1.0001)), add(scl(a, rad * cos(ang)), scl(b, rad * sin(ang))));
        *dir = neg(w);
        *nl_out = neg(w);
        return;
    }
    int l = 0;
    while (l < s->nl - 1 && s->light_cdf[l + 1] <= target) l++;
    const orc_light* L = &s->lights[l];
    v3 c = V(L->corner[0], L->corner[1], L->corner[2]);
    v3 eu = V(L->edge_u[0], L->edge_u[1], L->edge_u[2]);
    v3 ev = V(L->edge_v[0], L->edge_v[1], L->edge_v[2]);
    double u1 = rng_at(r, 1), u2 = rng_at(r, 2);
    *pos = add(c, add(scl(eu, u1), scl(ev, u2)));
    v3 nl = normalized(cross(eu, ev));
    *nl_out = nl;
    double wgt = emit_direction(pr, nl, u1, u2, r, dir, NULL);
    double pl = (s->light_cdf[l + 1] - s->light_cdf[l]) / total;
    double area = norm(cross(eu, ev));
    double k = PI * area / (pl * (double)N) * wgt;
    *flux = V(L->emission[0] * k, L->emission[1] * k, L->emission[2] * k);
}

/* One photon path (ppm_probabilistic.cc:132-193).  Lambertian hits store a
 * deposit in out[] (up to cap_out, the count is returned whatever the cap);
 * with irr != NULL a hit on a BSSRDF surface arriving from outside also stores
 * an irradiance deposit (the photons of the subsurface pre-pass,
 * subsurface_integrator.cc:240-327, taken from this pass's own photons) in
 * irr[] (count in *n_irr), and the path goes on through the dielectric
 * boundary. */
static int trace_one_photon(const orc_scene* s, const orc_pass_params* P, const orc_proposal* pr,
                            int64_t idx, deposit* out, int cap_out, deposit* irr, int cap_irr, int* n_irr) {
    rng_t r;
    rng_init(&r, P->seed, KIND_PHOTON, (uint32_t)P->pass_index, (uint32_t)idx, P->scene_id);
    v3 pos, dir, flux, nl;
    emit_photon(s, pr, &r, P->photons, &pos, &dir, &flux, &nl);
    v3 o = add(pos, scl(nl, P->ray_eps));
    int nd = 0, ni = 0;
    for (int bounce = 0;; bounce++) {
        if (bounce >= P->max_bounces || vmax3(flux) <= 0.0) break;
        double t;
        int32_t tri = scene_intersect(s, o, dir, &t);
        if (tri < 0) break;
        int dd = PH_DIM_BOUNCE + 3 * bounce;
        double r0 = rng_at(&r, dd), r1 = rng_at(&r, dd + 1), r2 = rng_at(&r, dd + 2);
        const orc_bsdf* b = &s->bsdf[s->mat[tri]];
        v3 hp = add(o, scl(dir, t));
        v3 ng = tri_normal(s, tri);
        v3 on = dot(ng, dir) < 0.0 ? ng : neg(ng);
        if (b->type & ORC_EMITTER) break;
        if (irr && (b->type & ORC_BSSRDF) && dot(ng, dir) < 0.0) {
            if (ni < cap_irr) {
                deposit* d = &irr[ni];
                d->p[0] = (float)hp.x; d->p[1] = (float)hp.y; d->p[2] = (float)hp.z;
                d->f[0] = (float)flux.x; d->f[1] = (float)flux.y; d->f[2] = (float)flux.z;
                d->n[0] = (float)on.x; d->n[1] = (float)on.y; d->n[2] = (float)on.z;
                d->bounce = bounce;
            }
            ni++;
        }
        if (b->type & ORC_LAMBERTIAN) {
            if (nd < cap_out) {
                deposit* d = &out[nd];
                d->p[0] = (float)hp.x; d->p[1] = (float)hp.y; d->p[2] = (float)hp.z;
                d->f[0] = (float)flux.x; d->f[1] = (float)flux.y; d->f[2] = (float)flux.z;
                d->n[0] = (float)on.x; d->n[1] = (float)on.y; d->n[2] = (float)on.z;
                d->bounce = bounce;
            }
            nd++;
            v3 rho = refl_of(b);
            double prob = (rho.x + rho.y + rho.z) / 3.0;
            if (r0 < prob) {
                double pdf = 1.0;
                v3 nd_dir = bsdf_sample(b, dir, ng, r1, r2, &pdf);
                o = spawn(hp, ng, nd_dir, P->ray_eps);
                dir = nd_dir;
                flux = scl(mul(flux, rho), 1.0 / prob);
            } else {
                break;
            }
        } else {
            double pdf = 1.0;
            /* the unoriented normal, as the eye pass passes it (ppm_probabilistic.cc:338);
             * the photon pass's oriented normal (:184) would make every glass hit "into" */
            v3 nd_dir = bsdf_sample(b, dir, ng, r0, r1, &pdf);
            o = spawn(hp, ng, nd_dir, P->ray_eps);
            dir = nd_dir;
            flux = scl(mul(flux, refl_of(b)), 1.0 / pdf);
        }
    }
    if (n_irr) *n_irr = ni;
    return nd;
}

#define MAX_DEP 64
#define CACHE_MAIN 4
#define CACHE_IRR 2

/* All deposits of a pass in (photon index, bounce) order -- the order the
 * sort of the photon map leaves them in on the GPU -- as two malloc'd lists
 * (main: Lambertian stores; irr: BSSRDF-surface stores, only when want_irr).
 * Photons with more deposits than the per-photon cache are traced again (the
 * path is a function of its index). */
static int64_t trace_pass(const orc_scene* s, const orc_pass_params* P, const orc_proposal* pr, int want_irr,
                          deposit** main_out, int32_t** main_photon, deposit** irr_out, int32_t** irr_photon,
                          int64_t* n_irr_out) {
    int64_t N = P->photons;
    int32_t* cm = (int32_t*)calloc((size_t)(N + 1), sizeof(int32_t));
    int32_t* ci = (int32_t*)calloc((size_t)(N + 1), sizeof(int32_t));
    deposit* tm = (deposit*)malloc(sizeof(deposit) * (size_t)(N ? N : 1) * CACHE_MAIN);
    deposit* ti = want_irr ? (deposit*)malloc(sizeof(deposit) * (size_t)(N ? N : 1) * CACHE_IRR) : NULL;
#pragma omp parallel for schedule(dynamic, 1024)
    for (int64_t i = 0; i < N; i++) {
        deposit buf[MAX_DEP], ib[MAX_DEP];
        int ni = 0;
        int nd = trace_one_photon(s, P, pr, i, buf, MAX_DEP, want_irr ? ib : NULL, MAX_DEP, &ni);
        cm[i + 1] = nd;
        ci[i + 1] = ni;
        for (int k = 0; k < nd && k < CACHE_MAIN; k++) tm[i * CACHE_MAIN + k] = buf[k];
        for (int k = 0; k < ni && k < CACHE_IRR; k++) ti[i * CACHE_IRR + k] = ib[k];
    }
    int64_t* om = (int64_t*)malloc(sizeof(int64_t) * (size_t)(N + 1));
    int64_t* oi = (int64_t*)malloc(sizeof(int64_t) * (size_t)(N + 1));
    om[0] = 0;
    oi[0] = 0;
    for (int64_t i = 0; i < N; i++) {
        om[i + 1] = om[i] + (cm[i + 1] < MAX_DEP ? cm[i + 1] : MAX_DEP);
        oi[i + 1] = oi[i] + (ci[i + 1] < MAX_DEP ? ci[i + 1] : MAX_DEP);
    }
    deposit* dm = (deposit*)malloc(sizeof(deposit) * (size_t)(om[N] ? om[N] : 1));
    int32_t* pm = (int32_t*)malloc(sizeof(int32_t) * (size_t)(om[N] ? om[N] : 1));
    deposit* di = (deposit*)malloc(sizeof(deposit) * (size_t)(oi[N] ? oi[N] : 1));
    int32_t* pi = (int32_t*)malloc(sizeof(int32_t) * (size_t)(oi[N] ? oi[N] : 1));
#pragma omp parallel for schedule(dynamic, 1024)
    for (int64_t i = 0; i < N; i++) {
        int nd = cm[i + 1], ni = ci[i + 1];
        deposit buf[MAX_DEP], ib[MAX_DEP];
        const deposit* sm = &tm[i * CACHE_MAIN];
        const deposit* si = ti ? &ti[i * CACHE_IRR] : NULL;
        if (nd > CACHE_MAIN || ni > CACHE_IRR) {
            int dummy = 0;
            trace_one_photon(s, P, pr, i, buf, MAX_DEP, want_irr ? ib : NULL, MAX_DEP, &dummy);
            sm = buf;
            si = ib;
        }
        for (int k = 0; k < nd && k < MAX_DEP; k++) { dm[om[i] + k] = sm[k]; pm[om[i] + k] = (int32_t)i; }
        for (int k = 0; k < ni && k < MAX_DEP; k++) { di[oi[i] + k] = si[k]; pi[oi[i] + k] = (int32_t)i; }
    }
    int64_t total = om[N];
    *n_irr_out = oi[N];
    *main_out = dm;
    *main_photon = pm;
    *irr_out = di;
    *irr_photon = pi;
    free(om); free(oi); free(tm); free(ti); free(cm); free(ci);
    return total;
}

int64_t orc_trace_photons(const orc_scene* s, const orc_pass_params* P, float* pos, float* flux,
                          float* nrm, uint32_t* key, int64_t cap) {
    return orc_trace_photons_prop(s, P, NULL, pos, flux, nrm, key, cap);
}

static void copy_deposits(const deposit* d, const int32_t* ph, int64_t n, int64_t cap, float* pos, float* flux,
                          float* nrm, uint32_t* key) {
    for (int64_t j = 0; j < n && j < cap; j++) {
        memcpy(pos + 3 * j, d[j].p, 12);
        memcpy(flux + 3 * j, d[j].f, 12);
        memcpy(nrm + 3 * j, d[j].n, 12);
        if (key) key[j] = ((uint32_t)ph[j] << 6) | (uint32_t)d[j].bounce;
    }
}

int64_t orc_trace_photons_prop(const orc_scene* s, const orc_pass_params* P, const orc_proposal* pr,
                               float* pos, float* flux, float* nrm, uint32_t* key, int64_t cap) {
    deposit *dm, *di;
    int32_t *pm, *pi;
    int64_t ni;
    int64_t total = trace_pass(s, P, pr, 0, &dm, &pm, &di, &pi, &ni);
    copy_deposits(dm, pm, total, cap, pos, flux, nrm, key);
    free(dm); free(pm); free(di); free(pi);
    return total;
}

int64_t orc_trace_irradiance_photons(const orc_scene* s, const orc_pass_params* P, float* pos, float* flux,
                                     float* nrm, uint32_t* key, int64_t cap) {
    deposit *dm, *di;
    int32_t *pm, *pi;
    int64_t ni;
    trace_pass(s, P, NULL, 1, &dm, &pm, &di, &pi, &ni);
    copy_deposits(di, pi, ni, cap, pos, flux, nrm, key);
    free(dm); free(pm); free(di); free(pi);
    return ni;
}

/* ------------------------------------------------------------------------ */
/* Eye pass (ppm_probabilistic.cc:210-343)                                   */
/* ------------------------------------------------------------------------ */

/* perspective_camera.cc:42-46 with an orthonormal (U, V) basis and the field
 * of view in degrees; pixel (x, row) with row 0 at the top of the image. */
static void camera_ray(const orc_camera* c, double px, double py_up, v3* o, v3* d) {
    v3 D = normalized(V(c->direction[0], c->direction[1], c->direction[2]));
    v3 up = V(c->up[0], c->up[1], c->up[2]);
    v3 U = normalized(cross(D, up));
    v3 Vv = cross(U, D);
    double th = tan(c->fov_y_deg * 0.5 * PI / 180.0);
    double aspect = (double)c->width / (double)c->height;
    v3 dir = add(D, add(scl(U, th * aspect * (2.0 * px / c->width - 1.0)),
                        scl(Vv, th * (2.0 * py_up / c->height - 1.0))));
    *o = V(c->center[0], c->center[1], c->center[2]);
    *d = normalized(dir);
}

/* Outcomes of an eye path */
#define EYE_NONE 0   /* black environment, killed, or an emitter seen (*direct) */
#define EYE_HIT 1    /* hit point on a Lambertian surface */
#define EYE_SSS 2    /* transmitted into a BSSRDF body: the dipole estimate at *hpos */

/* ppm_probabilistic.cc:237-343 (executePathTracing, radiance) in the P-KZ
 * form.  EYE_HIT: *hw = throughput * rho / (pi * roulette).  EYE_SSS:
 * *hw = throughput * F_t / (P_t * roulette) (the transmitted branch of
 * :272-290, divided by the roulette) and *mat = the material. */
static int eye_path_rr(const orc_scene* s, const orc_camera* cam, const orc_pass_params* P, int64_t pix,
                       v3* hpos, v3* hnrm, v3* hw, v3* direct, int32_t* mat, int roulette_on);

static int eye_path(const orc_scene* s, const orc_camera* cam, const orc_pass_params* P,
                    int64_t pix, v3* hpos, v3* hnrm, v3* hw, v3* direct, int32_t* mat) {
    return eye_path_rr(s, cam, P, pix, hpos, hnrm, hw, direct, mat, 1);
}

/* roulette_on = 0: no eye Russian roulette (the Hachisuka method's eye pass,
 * progressive_photon_mapping.cc:289-343) */
static int eye_path_rr(const orc_scene* s, const orc_camera* cam, const orc_pass_params* P, int64_t pix,
                       v3* hpos, v3* hnrm, v3* hw, v3* direct, int32_t* mat, int roulette_on) {
    int W = cam->width, H = cam->height;
    int x = (int)(pix % W), row = (int)(pix / W);
    rng_t r;
    rng_init(&r, P->seed, KIND_EYE, (uint32_t)P->pass_index, (uint32_t)pix, P->scene_id);
    double px = x + rng_at(&r, 0);
    double py = (H - 1 - row) + rng_at(&r, 1);
    v3 o, d;
    camera_ray(cam, px, py, &o, &d);
    v3 thr = V(1, 1, 1);
    *direct = V(0, 0, 0);
    for (int bounce = 0;; bounce++) {
        double t;
        int32_t tri = bounce >= P->max_bounces ? -1 : scene_intersect(s, o, d, &t);
        if (tri < 0) {
            /* the environment, or black (ppm_probabilistic.cc:248-251) */
            if (s->env_w > 0) *direct = mul(thr, env_lookup(s, d));
            return EYE_NONE;
        }
        int dd = EYE_DIM_BOUNCE + 3 * bounce;
        double r0 = rng_at(&r, dd), r1 = rng_at(&r, dd + 1), r2 = rng_at(&r, dd + 2);
        const orc_bsdf* b = &s->bsdf[s->mat[tri]];
        v3 hp = add(o, scl(d, t));
        v3 ng = tri_normal(s, tri);
        if (b->type & ORC_EMITTER) {
            if (dot(d, ng) < 0.0)
                *direct = mul(thr, V(b->emission[0], b->emission[1], b->emission[2]));
            return EYE_NONE;
        }
        v3 rho = refl_of(b);
        double roulette = vmax3(rho);
        if (roulette_on && bounce >= 3) {
            if (r0 > roulette) return EYE_NONE;
        } else {
            roulette = 1.0;
        }
        v3 on = dot(ng, d) < 0.0 ? ng : neg(ng);
        if (b->type & ORC_LAMBERTIAN) {
            *hpos = hp;
            *hnrm = on;
            *hw = scl(mul(thr, rho), 1.0 / (PI * roulette));
            return EYE_HIT;
        }
        double pdf = 1.0;
        v3 nd = bsdf_sample(b, d, ng, r1, r2, &pdf);
        if ((b->type & ORC_BSSRDF) && mat && ((dot(nd, ng) < 0.0) == (dot(d, ng) < 0.0))) {
            *hpos = hp;
            *hnrm = on;
            *hw = scl(thr, 1.0 / (pdf * roulette));
            *mat = s->mat[tri];
            return EYE_SSS;
        }
        thr = scl(mul(thr, rho), 1.0 / (pdf * roulette));
        o = spawn(hp, ng, nd, P->ray_eps);
        d = nd;
    }
}

void orc_eye_pass(const orc_scene* s, const orc_camera* cam, const orc_pass_params* P,
                  int32_t* has_hit, float* hit_pos, float* hit_nrm, float* hit_w,
                  double* direct_rgb) {
    int64_t np = (int64_t)cam->width * cam->height;
#pragma omp parallel for schedule(dynamic, 256)
    for (int64_t i = 0; i < np; i++) {
        v3 hp = V(0, 0, 0), hn = V(0, 0, 0), hw = V(0, 0, 0), dr;
        int h = eye_path(s, cam, P, i, &hp, &hn, &hw, &dr, NULL) == EYE_HIT;
        has_hit[i] = h;
        if (!h) hp = hn = hw = V(0, 0, 0);
        hit_pos[3 * i] = (float)hp.x; hit_pos[3 * i + 1] = (float)hp.y; hit_pos[3 * i + 2] = (float)hp.z;
        hit_nrm[3 * i] = (float)hn.x; hit_nrm[3 * i + 1] = (float)hn.y; hit_nrm[3 * i + 2] = (float)hn.z;
        hit_w[3 * i] = (float)hw.x; hit_w[3 * i + 1] = (float)hw.y; hit_w[3 * i + 2] = (float)hw.z;
        direct_rgb[3 * i] = dr.x; direct_rgb[3 * i + 1] = dr.y; direct_rgb[3 * i + 2] = dr.z;
    }
}

/* ------------------------------------------------------------------------ */
/* Gather (P-KZ form of ppm_probabilistic.cc:300-334)                        */
/* ------------------------------------------------------------------------ */

#define CONE_K 1.1

static void gather_one(const float* ph_pos, const float* ph_flux, const float* ph_nrm,
                       const int32_t* sorted, const int32_t* cs, const int32_t* ce, int32_t S,
                       double scale, const double bmin[3], double r, v3 x, v3 nx, v3 w,
                       double* out, int64_t* cnt) {
    int32_t lo[3], hi[3];
    double xs[3] = {x.x, x.y, x.z};
    for (int d = 0; d < 3; d++) {
        /* [x - r, x + r] has the width of one cell (scale = 1 / 2r): it meets at
         * most two cells per axis; a third only appears through FP64 rounding at
         * an exact cell boundary and is dropped (as on the GPU) */
        lo[d] = cell_coord(xs[d] - r, bmin[d], scale);
        hi[d] = cell_coord(xs[d] + r, bmin[d], scale);
        if (hi[d] > lo[d] + 1) hi[d] = lo[d] + 1;
    }
    uint32_t seen[8];
    int ns = 0;
    double r2 = r * r;
    v3 acc = V(0, 0, 0);
    int64_t nb = 0, nc = 0, ni = 0;
    for (int32_t iz = lo[2]; iz <= hi[2]; iz++)
        for (int32_t iy = lo[1]; iy <= hi[1]; iy++)
            for (int32_t ix = lo[0]; ix <= hi[0]; ix++) {
                uint32_t h = orc_hash(ix, iy, iz, S);
                int dup = 0;
                for (int k = 0; k < ns; k++) if (seen[k] == h) dup = 1;
                if (dup) continue;
                if (ns < 8) seen[ns++] = h;
                nb++;
                for (int32_t j = cs[h]; j < ce[h]; j++) {
                    int32_t pi = sorted[j];
                    nc++;
                    v3 p = V(ph_pos[3 * pi], ph_pos[3 * pi + 1], ph_pos[3 * pi + 2]);
                    v3 df = sub(x, p);
                    double d2 = dot(df, df);
                    if (d2 > r2) continue;
                    v3 np = V(ph_nrm[3 * pi], ph_nrm[3 * pi + 1], ph_nrm[3 * pi + 2]);
                    if (dot(np, nx) <= 0.0) continue;
                    ni++;
                    double wt = 1.0 - sqrt(d2) / (CONE_K * r);
                    acc = add(acc, scl(V(ph_flux[3 * pi], ph_flux[3 * pi + 1], ph_flux[3 * pi + 2]), wt));
                }
            }
    double norm_ = PI * r2 * (1.0 - 2.0 / (3.0 * CONE_K));
    out[0] = w.x * acc.x / norm_;
    out[1] = w.y * acc.y / norm_;
    out[2] = w.z * acc.z / norm_;
    if (cnt) { cnt[0] = nb; cnt[1] = nc; cnt[2] = ni; }
}

void orc_gather(const float* ph_pos, const float* ph_flux, const float* ph_nrm, int64_t n_ph,
                const float* hit_pos, const float* hit_nrm, const float* hit_w, int64_t n_hit,
                double radius, int32_t S, const double bmin[3], double* out, int64_t* counters) {
    double scale = 1.0 / (2.0 * radius);
    int32_t* sorted = (int32_t*)malloc(sizeof(int32_t) * (size_t)(n_ph ? n_ph : 1));
    int32_t* cs = (int32_t*)malloc(sizeof(int32_t) * (size_t)S);
    int32_t* ce = (int32_t*)malloc(sizeof(int32_t) * (size_t)S);
    orc_hashgrid_points(ph_pos, n_ph, S, scale, bmin, sorted, cs, ce);
    int64_t tb = 0, tc = 0, ti = 0;
#pragma omp parallel for schedule(dynamic, 256) reduction(+ : tb, tc, ti)
    for (int64_t i = 0; i < n_hit; i++) {
        v3 x = V(hit_pos[3 * i], hit_pos[3 * i + 1], hit_pos[3 * i + 2]);
        v3 nx = V(hit_nrm[3 * i], hit_nrm[3 * i + 1], hit_nrm[3 * i + 2]);
        v3 w = V(hit_w[3 * i], hit_w[3 * i + 1], hit_w[3 * i + 2]);
        int64_t c[3];
        gather_one(ph_pos, ph_flux, ph_nrm, sorted, cs, ce, S, scale, bmin, radius, x, nx, w,
                   out + 3 * i, c);
        tb += c[0]; tc += c[1]; ti += c[2];
    }
    if (counters) { counters[0] = tb; counters[1] = tc; counters[2] = ti; }
    free(sorted); free(cs); free(ce);
}

/* ------------------------------------------------------------------------ */
/* Pass loop (ppm_probabilistic.cc:24-121)                                   */
/* ------------------------------------------------------------------------ */

/* (1 - F_dr) / pi * area * sum_j R_d(|x - p_j|^2) E_j for one point x and
 * material b (subsurface_integrator.cc:147-167 without the octree cut, and
 * :234-238), FP64 with the true exponential */
static void dipole_point(const orc_bsdf* b, const float* pts, const double* E, int64_t n_pts, double area, v3 x,
                         double out[3]) {
    double sa[3] = {b->sigma_a[0], b->sigma_a[1], b->sigma_a[2]};
    double ss[3] = {b->sigmap_s[0], b->sigmap_s[1], b->sigmap_s[2]};
    double f = (1.0 - orc_dipole_fdr(b->eta)) / PI * area;
    double acc[3] = {0.0, 0.0, 0.0};
    for (int64_t j = 0; j < n_pts; j++) {
        double dx = x.x - pts[3 * j], dy = x.y - pts[3 * j + 1], dz = x.z - pts[3 * j + 2];
        double rd[3];
        orc_dipole(sa, ss, b->eta, dx * dx + dy * dy + dz * dz, 0, rd);
        for (int c = 0; c < 3; c++) acc[c] += rd[c] * E[3 * j + c];
    }
    for (int c = 0; c < 3; c++) out[c] = f * acc[c];
}

int32_t orc_render_ex(const orc_scene* s, const orc_camera* cam, int64_t photons, int32_t n_passes,
                      int32_t first_pass, double r1, double alpha, double ray_eps, uint64_t seed,
                      uint32_t scene_id, int32_t hash_size, const orc_render_opts* opts, double* image,
                      int64_t* stats) {
    int64_t np = (int64_t)cam->width * cam->height;
    const orc_proposal* pr = opts ? opts->prop : NULL;
    const int sss = opts && opts->sss_n > 0 && opts->sss_pts;
    double* radii = (double*)malloc(sizeof(double) * (size_t)(first_pass + n_passes));
    orc_radius_schedule(r1, alpha, first_pass + n_passes, 0, radii);
    double bmin[3] = {s->bmin.x - r1, s->bmin.y - r1, s->bmin.z - r1};
    int32_t* kind = (int32_t*)malloc(sizeof(int32_t) * (size_t)np);
    int32_t* smat = (int32_t*)malloc(sizeof(int32_t) * (size_t)np);
    float* hp = (float*)malloc(sizeof(float) * 3 * (size_t)np);
    float* hn = (float*)malloc(sizeof(float) * 3 * (size_t)np);
    float* hw = (float*)malloc(sizeof(float) * 3 * (size_t)np);
    double* dr = (double*)malloc(sizeof(double) * 3 * (size_t)np);
    double* g = (double*)malloc(sizeof(double) * 3 * (size_t)np);
    int64_t* hidx = (int64_t*)malloc(sizeof(int64_t) * (size_t)np);
    float* cp = (float*)malloc(sizeof(float) * 3 * (size_t)np);
    float* cn = (float*)malloc(sizeof(float) * 3 * (size_t)np);
    float* cw = (float*)malloc(sizeof(float) * 3 * (size_t)np);
    double* sum = (double*)calloc(3 * (size_t)np, sizeof(double));
    double* E = sss ? (double*)malloc(sizeof(double) * 3 * (size_t)opts->sss_n) : NULL;
    float* ptw = NULL;
    if (sss) {
        ptw = (float*)malloc(sizeof(float) * 3 * (size_t)opts->sss_n);
        for (int64_t j = 0; j < 3 * opts->sss_n; j++) ptw[j] = (float)(1.0 / PI);
    }
    int64_t tot_dep = 0, tot_hit = 0, tot_irr = 0, tot_sss = 0;
    int32_t rc = 0;
    for (int32_t k = 0; k < n_passes; k++) {
        int32_t pass = first_pass + k;
        orc_pass_params P;
        memset(&P, 0, sizeof(P));
        P.photons = photons;
        P.pass_index = pass;
        P.max_bounces = 64;
        P.radius = radii[pass - 1];
        P.ray_eps = ray_eps;
        P.seed = seed;
        P.scene_id = scene_id;
        deposit *dm, *di;
        int32_t *pm, *pi;
        int64_t ni;
        int64_t nd = trace_pass(s, &P, pr, sss, &dm, &pm, &di, &pi, &ni);
        tot_dep += nd;
        tot_irr += ni;
        float* pp = (float*)malloc(sizeof(float) * 3 * (size_t)(nd ? nd : 1));
        float* pf = (float*)malloc(sizeof(float) * 3 * (size_t)(nd ? nd : 1));
        float* pn = (float*)malloc(sizeof(float) * 3 * (size_t)(nd ? nd : 1));
        copy_deposits(dm, pm, nd, nd, pp, pf, pn, NULL);
        if (sss) {
            /* irradiance at the surface points: the fixed-radius gather of the
             * irradiance photons with weight 1 / pi (subsurface_integrator.cc:329-365) */
            if (ni > 0) {
                float* ip = (float*)malloc(sizeof(float) * 3 * (size_t)ni);
                float* iff = (float*)malloc(sizeof(float) * 3 * (size_t)ni);
                float* in_ = (float*)malloc(sizeof(float) * 3 * (size_t)ni);
                copy_deposits(di, pi, ni, ni, ip, iff, in_, NULL);
                int32_t Si = hash_size > 0 ? hash_size : (int32_t)ni;
                orc_gather(ip, iff, in_, ni, opts->sss_pts, opts->sss_nrm, ptw, opts->sss_n, P.radius, Si, bmin, E,
                           NULL);
                free(ip); free(iff); free(in_);
            } else {
                memset(E, 0, sizeof(double) * 3 * (size_t)opts->sss_n);
            }
        }
        free(dm); free(pm); free(di); free(pi);
#pragma omp parallel for schedule(dynamic, 256)
        for (int64_t i = 0; i < np; i++) {
            v3 a = V(0, 0, 0), n_ = V(0, 0, 0), w = V(0, 0, 0), d;
            int32_t m = -1;
            kind[i] = eye_path(s, cam, &P, i, &a, &n_, &w, &d, sss ? &m : NULL);
            smat[i] = m;
            hp[3 * i] = (float)a.x; hp[3 * i + 1] = (float)a.y; hp[3 * i + 2] = (float)a.z;
            hn[3 * i] = (float)n_.x; hn[3 * i + 1] = (float)n_.y; hn[3 * i + 2] = (float)n_.z;
            hw[3 * i] = (float)w.x; hw[3 * i + 1] = (float)w.y; hw[3 * i + 2] = (float)w.z;
            dr[3 * i] = d.x; dr[3 * i + 1] = d.y; dr[3 * i + 2] = d.z;
        }
        int64_t nh = 0;
        for (int64_t i = 0; i < np; i++)
            if (kind[i] == EYE_HIT) {
                hidx[nh] = i;
                memcpy(cp + 3 * nh, hp + 3 * i, 12);
                memcpy(cn + 3 * nh, hn + 3 * i, 12);
                memcpy(cw + 3 * nh, hw + 3 * i, 12);
                nh++;
            }
        tot_hit += nh;
        int32_t S = hash_size > 0 ? hash_size : (int32_t)(nd > 0 ? nd : 1);
        orc_gather(pp, pf, pn, nd, cp, cn, cw, nh, P.radius, S, bmin, g, NULL);
        for (int64_t i = 0; i < np; i++)
            for (int c = 0; c < 3; c++) sum[3 * i + c] += dr[3 * i + c];
        for (int64_t j = 0; j < nh; j++)
            for (int c = 0; c < 3; c++) sum[3 * hidx[j] + c] += g[3 * j + c];
        if (sss) {
            int64_t ns = 0;
#pragma omp parallel for schedule(dynamic, 16) reduction(+ : ns)
            for (int64_t i = 0; i < np; i++) {
                if (kind[i] != EYE_SSS) continue;
                ns++;
                double L[3];
                dipole_point(&s->bsdf[smat[i]], opts->sss_pts, E, opts->sss_n, opts->sss_area,
                             V(hp[3 * i], hp[3 * i + 1], hp[3 * i + 2]), L);
                for (int c = 0; c < 3; c++) sum[3 * i + c] += (double)hw[3 * i + c] * L[c];
            }
            tot_sss += ns;
        }
        free(pp); free(pf); free(pn);
    }
    for (int64_t i = 0; i < 3 * np; i++) image[i] = sum[i] / (double)n_passes;
    if (stats) { stats[0] = tot_dep; stats[1] = tot_hit; stats[2] = tot_irr; stats[3] = tot_sss; }
    free(radii); free(kind); free(smat); free(hp); free(hn); free(hw);
    free(dr); free(g); free(hidx); free(cp); free(cn); free(cw); free(sum); free(E); free(ptw);
    return rc;
}

int32_t orc_render(const orc_scene* s, const orc_camera* cam, int64_t photons, int32_t n_passes,
                   int32_t first_pass, double r1, double alpha, double ray_eps, uint64_t seed,
                   uint32_t scene_id, int32_t hash_size, double* image, int64_t* stats) {
    int64_t st[4];
    int32_t rc = orc_render_ex(s, cam, photons, n_passes, first_pass, r1, alpha, ray_eps, seed, scene_id,
                               hash_size, NULL, image, st);
    if (stats) { stats[0] = st[0]; stats[1] = st[1]; }
    return rc;
}

/* ------------------------------------------------------------------------ */
/* Independent FP64 path tracer with next-event estimation                   */
/* ------------------------------------------------------------------------ */

#define KIND_PT 3u
/* dims: 0-1 jitter; bounce b uses 2+6b .. 7+6b: roulette, direction (2),
 * light pick, light point (2) */
#define PT_DIM_BOUNCE 2

/* L at the camera through each pixel: emission seen directly or through
 * specular chains, and at every diffuse vertex the direct light by sampling
 * a point on a light (shadow ray) plus the indirect light by continuing the
 * path with a cosine-distributed direction.  An emitter reached from a
 * diffuse vertex adds nothing (its light was counted by the light sample).
 * Consistent with the photon mapping estimate in the limit of zero radius;
 * shares only the scene, intersection and the pinned BSDF sampling with it.
 * (The reference's own PathTracing, path_tracing.cc:136-198, sees only the
 * environment map, so it cannot render an area-lit scene.) */
static v3 pt_path(const orc_scene* s, const orc_camera* cam, int64_t pix, int32_t sample, int32_t max_bounces,
                  double ray_eps, uint64_t seed, uint32_t scene_id) {
    int W = cam->width, H = cam->height;
    int x = (int)(pix % W), row = (int)(pix / W);
    rng_t r;
    rng_init(&r, seed, KIND_PT, (uint32_t)sample, (uint32_t)pix, scene_id);
    v3 o, d;
    camera_ray(cam, x + rng_at(&r, 0), (H - 1 - row) + rng_at(&r, 1), &o, &d);
    v3 thr = V(1, 1, 1), L = V(0, 0, 0);
    int specular_prev = 1;
    for (int bounce = 0; bounce < max_bounces; bounce++) {
        double t;
        int32_t tri = scene_intersect(s, o, d, &t);
        if (tri < 0) {
            /* the environment is not sampled at diffuse vertices: count it on every miss */
            if (s->env_w > 0) L = add(L, mul(thr, env_lookup(s, d)));
            break;
        }
        const orc_bsdf* b = &s->bsdf[s->mat[tri]];
        v3 hp = add(o, scl(d, t));
        v3 ng = tri_normal(s, tri);
        if (b->type & ORC_EMITTER) {
            if (specular_prev && dot(d, ng) < 0.0)
                L = add(L, mul(thr, V(b->emission[0], b->emission[1], b->emission[2])));
            break;
        }
        int dd = PT_DIM_BOUNCE + 6 * bounce;
        v3 rho = refl_of(b);
        if (bounce >= 3) {
            double q = vmax3(rho);
            if (rng_at(&r, dd) >= q) break;
            thr = scl(thr, 1.0 / q);
        }
        v3 on = dot(ng, d) < 0.0 ? ng : neg(ng);
        if (b->type & ORC_LAMBERTIAN) {
            if (s->nl > 0) {
                double target = rng_at(&r, dd + 3) * s->light_total;
                int l = 0;
                while (l < s->nl - 1 && s->light_cdf[l + 1] <= target) l++;
                const orc_light* Lt = &s->lights[l];
                v3 c = V(Lt->corner[0], Lt->corner[1], Lt->corner[2]);
                v3 eu = V(Lt->edge_u[0], Lt->edge_u[1], Lt->edge_u[2]);
                v3 ev = V(Lt->edge_v[0], Lt->edge_v[1], Lt->edge_v[2]);
                v3 y = add(c, add(scl(eu, rng_at(&r, dd + 4)), scl(ev, rng_at(&r, dd + 5))));
                v3 cr = cross(eu, ev);
                double area = norm(cr);
                v3 nl = scl(cr, 1.0 / area);
                double pl = (s->light_cdf[l + 1] - s->light_cdf[l]) / s->light_total;
                v3 w = sub(y, hp);
                double dist2 = dot(w, w), dist = sqrt(dist2);
                w = scl(w, 1.0 / dist);
                double cx = dot(on, w), cl = -dot(nl, w);
                if (cx > 0.0 && cl > 0.0) {
                    v3 so = spawn(hp, ng, w, ray_eps);
                    v3 ws = sub(y, so);
                    double ds = norm(ws);
                    double ts;
                    int32_t occ = scene_intersect(s, so, scl(ws, 1.0 / ds), &ts);
                    int visible = occ < 0 || ts >= ds * (1.0 - 1e-9) - 1e-9 ||
                                  (s->bsdf[s->mat[occ]].type & ORC_EMITTER);
                    if (visible) {
                        double gterm = cx * cl / dist2 * area / pl / PI;
                        L = add(L, scl(mul(mul(thr, rho), V(Lt->emission[0], Lt->emission[1], Lt->emission[2])),
                                       gterm));
                    }
                }
            }
            v3 nd = on_hemisphere(on, rng_at(&r, dd + 1), rng_at(&r, dd + 2));
            thr = mul(thr, rho);
            o = spawn(hp, ng, nd, ray_eps);
            d = nd;
            specular_prev = 0;
        } else {
            double pdf = 1.0;
            v3 nd = bsdf_sample(b, d, ng, rng_at(&r, dd + 1), rng_at(&r, dd + 2), &pdf);
            thr = scl(mul(thr, rho), 1.0 / pdf);
            o = spawn(hp, ng, nd, ray_eps);
            d = nd;
            specular_prev = 1;
        }
    }
    return L;
}

void orc_pathtrace(const orc_scene* s, const orc_camera* cam, int32_t spp, int32_t first_sample,
                   int32_t max_bounces, double ray_eps, uint64_t seed, uint32_t scene_id, double* image) {
    int64_t np = (int64_t)cam->width * cam->height;
#pragma omp parallel for schedule(dynamic, 16)
    for (int64_t i = 0; i < np; i++) {
        v3 acc = V(0, 0, 0);
        for (int32_t k = 0; k < spp; k++)
            acc = add(acc, pt_path(s, cam, i, first_sample + k, max_bounces, ray_eps, seed, scene_id));
        image[3 * i] = acc.x / spp;
        image[3 * i + 1] = acc.y / spp;
        image[3 * i + 2] = acc.z / spp;
    }
}

/* ------------------------------------------------------------------------ */
/* Progressive photon mapping with per-point radii                           */
/* (sources/progressive_photon_mapping.cc:22-287)                            */
/* ------------------------------------------------------------------------ */

/* A render point per pixel keeps its radius^2, photon count and flux across
 * passes.  Per pass: eye paths without roulette to their first Lambertian
 * surface (:289-343; what they see of emitters and the environment is the
 * pixel's emission); the initial radius from the hit points' extent
 * (:116-125); photons; then each render point takes the photons of its box
 * [x - R, x + R] one at a time -- the deduplicated buckets of its cells in z,
 * y, x order, each in photon order -- through the update of :244-258
 * (g = (n alpha + alpha) / (n alpha + 1)).  The reference's threads set the
 * order of the updates; this is one fixed order.  Image (:83-86):
 * (emission + tau / (pi r^2)) / passes. */
int32_t orc_render_hachisuka(const orc_scene* s, const orc_camera* cam, int64_t photons, int32_t n_passes,
                             double alpha, double ray_eps, uint64_t seed, uint32_t scene_id, int32_t hash_size,
                             double* image, int64_t* stats) {
    int64_t np = (int64_t)cam->width * cam->height;
    double* r2 = (double*)calloc((size_t)np, sizeof(double));
    int64_t* cnt = (int64_t*)calloc((size_t)np, sizeof(int64_t));
    double* tau = (double*)calloc(3 * (size_t)np, sizeof(double));
    double* emis = (double*)calloc(3 * (size_t)np, sizeof(double));
    int32_t* kind = (int32_t*)malloc(sizeof(int32_t) * (size_t)np);
    v3* xp = (v3*)malloc(sizeof(v3) * (size_t)np);
    v3* xn = (v3*)malloc(sizeof(v3) * (size_t)np);
    v3* xw = (v3*)malloc(sizeof(v3) * (size_t)np);
    int64_t tot_acc = 0;
    for (int32_t k = 0; k < n_passes; k++) {
        orc_pass_params P;
        memset(&P, 0, sizeof(P));
        P.photons = photons;
        P.pass_index = k + 1;
        P.max_bounces = 64;
        P.ray_eps = ray_eps;
        P.seed = seed;
        P.scene_id = scene_id;
        /* hit points: FP32 positions, as the GPU stores them */
        double lo[3] = {INFTY, INFTY, INFTY}, hi[3] = {-INFTY, -INFTY, -INFTY};
        int64_t nh = 0;
        for (int64_t i = 0; i < np; i++) {
            v3 a = V(0, 0, 0), n_ = V(0, 0, 0), w = V(0, 0, 0), d;
            kind[i] = eye_path_rr(s, cam, &P, i, &a, &n_, &w, &d, NULL, 0);
            emis[3 * i] += d.x;
            emis[3 * i + 1] += d.y;
            emis[3 * i + 2] += d.z;
            if (kind[i] != EYE_HIT) continue;
            xp[i] = V((float)a.x, (float)a.y, (float)a.z);
            xn[i] = V((float)n_.x, (float)n_.y, (float)n_.z);
            xw[i] = V((float)w.x, (float)w.y, (float)w.z);
            const double c[3] = {xp[i].x, xp[i].y, xp[i].z};
            for (int d2 = 0; d2 < 3; d2++) {
                lo[d2] = fmin(lo[d2], c[d2]);
                hi[d2] = fmax(hi[d2], c[d2]);
            }
            nh++;
        }
        deposit *dm, *di;
        int32_t *pm, *pi;
        int64_t ni;
        int64_t nd = trace_pass(s, &P, NULL, 0, &dm, &pm, &di, &pi, &ni);
        if (nh > 0) {
            double ext = ((hi[0] - lo[0]) + (hi[1] - lo[1]) + (hi[2] - lo[2])) / 3.0;
            double irad = ext / ((cam->width + cam->height) / 2.0) * 8.0;
            double bm[3] = {s->bmin.x - irad, s->bmin.y - irad, s->bmin.z - irad};
            double scale = 1.0 / (2.0 * irad);
            int32_t S = hash_size > 0 ? hash_size : (int32_t)(nd > 0 ? nd : 1);
            float* pp = (float*)malloc(sizeof(float) * 3 * (size_t)(nd ? nd : 1));
            float* pf = (float*)malloc(sizeof(float) * 3 * (size_t)(nd ? nd : 1));
            float* pn = (float*)malloc(sizeof(float) * 3 * (size_t)(nd ? nd : 1));
            copy_deposits(dm, pm, nd, nd, pp, pf, pn, NULL);
            int32_t* sorted = (int32_t*)malloc(sizeof(int32_t) * (size_t)(nd ? nd : 1));
            int32_t* cs = (int32_t*)malloc(sizeof(int32_t) * (size_t)S);
            int32_t* ce = (int32_t*)malloc(sizeof(int32_t) * (size_t)S);
            orc_hashgrid_points(pp, nd, S, scale, bm, sorted, cs, ce);
#pragma omp parallel for schedule(dynamic, 64) reduction(+ : tot_acc)
            for (int64_t i = 0; i < np; i++) {
                if (kind[i] != EYE_HIT) continue;
                double rr = r2[i];
                int64_t n = cnt[i];
                v3 t = V(tau[3 * i], tau[3 * i + 1], tau[3 * i + 2]);
                if (n == 0) {
                    rr = irad * irad;
                    t = V(0, 0, 0);
                }
                double xs[3] = {xp[i].x, xp[i].y, xp[i].z};
                int32_t cl[3], ch[3];
                for (int d = 0; d < 3; d++) {
                    cl[d] = cell_coord(xs[d] - irad, bm[d], scale);
                    ch[d] = cell_coord(xs[d] + irad, bm[d], scale);
                    if (ch[d] > cl[d] + 1) ch[d] = cl[d] + 1;
                }
                uint32_t seen[8];
                int ns = 0;
                for (int32_t iz = cl[2]; iz <= ch[2]; iz++)
                    for (int32_t iy = cl[1]; iy <= ch[1]; iy++)
                        for (int32_t ix = cl[0]; ix <= ch[0]; ix++) {
                            uint32_t h = orc_hash(ix, iy, iz, S);
                            int dup = 0;
                            for (int q = 0; q < ns; q++) dup |= seen[q] == h;
                            if (dup) continue;
                            seen[ns++] = h;
                            for (int32_t j = cs[h]; j < ce[h]; j++) {
                                int32_t id = sorted[j];
                                v3 pnr = V(pn[3 * id], pn[3 * id + 1], pn[3 * id + 2]);
                                if (!(dot(pnr, xn[i]) > 0.0)) continue;
                                v3 df = sub(V(xs[0], xs[1], xs[2]), V(pp[3 * id], pp[3 * id + 1], pp[3 * id + 2]));
                                if (!(dot(df, df) <= rr)) continue;
                                double na = (double)n * alpha;
                                double g = (na + alpha) / (na + 1.0);
                                rr *= g;
                                n++;
                                t = scl(add(t, mul(xw[i], V(pf[3 * id], pf[3 * id + 1], pf[3 * id + 2]))), g);
                                tot_acc++;
                            }
                        }
                r2[i] = rr;
                cnt[i] = n;
                tau[3 * i] = t.x;
                tau[3 * i + 1] = t.y;
                tau[3 * i + 2] = t.z;
            }
            free(pp); free(pf); free(pn); free(sorted); free(cs); free(ce);
        }
        free(dm); free(pm); free(di); free(pi);
    }
    for (int64_t i = 0; i < np; i++)
        for (int c = 0; c < 3; c++) {
            double tt = cnt[i] > 0 ? tau[3 * i + c] / (PI * r2[i]) : 0.0;
            image[3 * i + c] = (emis[3 * i + c] + tt) / n_passes;
        }
    if (stats) stats[0] = tot_acc;
    free(r2); free(cnt); free(tau); free(emis); free(kind); free(xp); free(xn); free(xw);
    return 0;
}

int32_t orc_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* ------------------------------------------------------------------------ */
/* Exported single-sample entry points (pinned against the compiled          */
/* reference in tests)                                                       */
/* ------------------------------------------------------------------------ */

void orc_on_hemisphere(const double n[3], double r1, double r2, double out[3]) {
    v3 d = on_hemisphere(V(n[0], n[1], n[2]), r1, r2);
    out[0] = d.x; out[1] = d.y; out[2] = d.z;
}

int32_t orc_total_reflection(int32_t into, const double in[3], const double n[3], const double on[3],
                             double refl[3], double refr[3], double* fre, double* ftr) {
    v3 rd, td;
    int r = total_reflection(into, V(in[0], in[1], in[2]), V(n[0], n[1], n[2]),
                             V(on[0], on[1], on[2]), IOR_OBJECT, &rd, &td, fre, ftr);
    refl[0] = rd.x; refl[1] = rd.y; refl[2] = rd.z;
    refr[0] = td.x; refr[1] = td.y; refr[2] = td.z;
    return r;
}

void orc_bsdf_sample(int32_t type, const double rho[3], const double in[3], const double n[3],
                     double r1, double r2, double out[3], double* pdf) {
    orc_bsdf b;
    memset(&b, 0, sizeof(b));
    b.type = type;
    b.reflectance[0] = (float)rho[0]; b.reflectance[1] = (float)rho[1]; b.reflectance[2] = (float)rho[2];
    double p = 1.0;
    v3 d = bsdf_sample(&b, V(in[0], in[1], in[2]), V(n[0], n[1], n[2]), r1, r2, &p);
    out[0] = d.x; out[1] = d.y; out[2] = d.z;
    *pdf = p;
}

void orc_camera_ray(const orc_camera* c, double px, double py_up, double o[3], double d[3]) {
    v3 oo, dd;
    camera_ray(c, px, py_up, &oo, &dd);
    o[0] = oo.x; o[1] = oo.y; o[2] = oo.z;
    d[0] = dd.x; d[1] = dd.y; d[2] = dd.z;
}

int32_t orc_tri_intersect(const double tri[9], const double o[3], const double d[3], double* t) {
    orc_scene s;
    memset(&s, 0, sizeof(s));
    v3 p0 = V(tri[0], tri[1], tri[2]), p1 = V(tri[3], tri[4], tri[5]), p2 = V(tri[6], tri[7], tri[8]);
    s.n = 1; s.p0 = &p0; s.p1 = &p1; s.p2 = &p2;
    double tt = tri_intersect(&s, 0, V(o[0], o[1], o[2]), V(d[0], d[1], d[2]));
    *t = tt;
    return tt > 0.0 ? 1 : 0;
}

void orc_emission(const orc_proposal* pr, const orc_pass_params* P, int64_t idx, const double nl[3],
                  double out_uv[2], double out_dir[3], double* weight, uint32_t* cell) {
    rng_t r;
    rng_init(&r, P->seed, KIND_PHOTON, (uint32_t)P->pass_index, (uint32_t)idx, P->scene_id);
    double u1 = rng_at(&r, 1), u2 = rng_at(&r, 2);
    v3 d;
    uint32_t c = 0;
    if (!pr || pr->side == 0) {
        double r3 = rng_at(&r, 3), r4 = rng_at(&r, 4);
        d = on_hemisphere(V(nl[0], nl[1], nl[2]), r3, r4);
        double ct = sqrt(1.0 - r4);
        int kz = (int)((1.0 - ct) * 32.0), kp = (int)(r3 * 32.0);
        c = (uint32_t)((kz > 31 ? 31 : kz) * 32 + (kp > 31 ? 31 : kp));
        *weight = 1.0;
    } else {
        *weight = emit_direction(pr, V(nl[0], nl[1], nl[2]), u1, u2, &r, &d, &c);
    }
    out_uv[0] = u1;
    out_uv[1] = u2;
    out_dir[0] = d.x;
    out_dir[1] = d.y;
    out_dir[2] = d.z;
    *cell = c;
}

/* ------------------------------------------------------------------------ */
/* Dipole BSSRDF (sources/bssrdf.cc:33-104) and its surface sum              */
/* (sources/subsurface_integrator.cc:147-167, 234-238)                      */
/* ------------------------------------------------------------------------ */

/* Schraudolph's exponential as the reference writes it (vector3d.cc:11-16):
 * only the high word of the double is set */
static double schraudolph_exp(double y) {
    int32_t hi = (int32_t)(1512775 * y + 1072632447);
    uint64_t bits = ((uint64_t)(uint32_t)hi) << 32;
    double d;
    memcpy(&d, &bits, sizeof d);
    return d;
}

double orc_dipole_fdr(double eta) {
    double e2 = eta * eta, e3 = e2 * eta;
    if (eta >= 1.0) return (-1.439 + 0.7099 * eta + 0.6681 * e2 + 0.0636 * e3) / e2;
    return (-0.4399 * e3 + 0.7099 * e2 - 0.3319 * eta + 0.0636) / e3;
}

/* R_d(d2) per channel, clamped at 0; fast = 1 uses the reference's exp */
void orc_dipole(const double sigma_a[3], const double sigmap_s[3], double eta, double d2, int32_t fast,
                double out[3]) {
    double fdr = orc_dipole_fdr(eta);
    double A = (1.0 + fdr) / (1.0 - fdr);
    for (int c = 0; c < 3; c++) {
        double st = sigma_a[c] + sigmap_s[c];
        double tr = sqrt(3.0 * sigma_a[c] * st);
        double ap = sigmap_s[c] / st;
        double zp = 1.0 / st;
        double zn = zp * (1.0 + (4.0 / 3.0) * A);
        double dp = sqrt(d2 + zp * zp), dn = sqrt(d2 + zn * zn);
        double ep = fast ? schraudolph_exp(-tr * dp) : exp(-tr * dp);
        double en = fast ? schraudolph_exp(-tr * dn) : exp(-tr * dn);
        double pos = zp * (dp * tr + 1.0) * ep;
        double neg = zn * (dn * tr + 1.0) * en;
        double r = (ap * (pos * dn * dn * dn - neg * dp * dp * dp)) / ((4.0 * PI) * tr * dp * dp * dp * dn * dn * dn);
        out[c] = r > 0.0 ? r : 0.0;
    }
}

/* out[i] = (1 - F_dr) / pi * area * sum_j R_d(|q_i - p_j|^2) E_j, FP64 brute force */
void orc_dipole_sum(const double sigma_a[3], const double sigmap_s[3], double eta, const float* pts,
                    const float* E, int64_t n_pts, double area, const float* q, int64_t n_q, double* out) {
    double f = (1.0 - orc_dipole_fdr(eta)) / PI * area;
#pragma omp parallel for schedule(dynamic, 16)
    for (int64_t i = 0; i < n_q; i++) {
        double acc[3] = {0.0, 0.0, 0.0};
        for (int64_t j = 0; j < n_pts; j++) {
            double dx = (double)q[3 * i] - pts[3 * j], dy = (double)q[3 * i + 1] - pts[3 * j + 1];
            double dz = (double)q[3 * i + 2] - pts[3 * j + 2];
            double r[3];
            orc_dipole(sigma_a, sigmap_s, eta, dx * dx + dy * dy + dz * dz, 0, r);
            for (int c = 0; c < 3; c++) acc[c] += r[c] * (double)E[3 * j + c];
        }
        for (int c = 0; c < 3; c++) out[3 * i + c] = f * acc[c];
    }
}

/* Octree over the surface points, FP64 (subsurface_integrator.cc:69-145):
 * a node's box is the tight box of its points, its children split that box
 * at the centre, child (x >= mid ? 4 : 0) + (y >= mid ? 2 : 0) + (z >= mid ? 1 : 0) */
typedef struct {
    double pos[3], irad[3], area, lo[3], hi[3];
    int32_t child[8];
    int32_t leaf;
} oct_node;

typedef struct {
    oct_node* n;
    int64_t len, cap;
} oct_tree;

static int32_t oct_new(oct_tree* t) {
    if (t->len == t->cap) {
        t->cap = t->cap ? 2 * t->cap : 1024;
        t->n = (oct_node*)realloc(t->n, sizeof(oct_node) * (size_t)t->cap);
    }
    memset(&t->n[t->len], 0, sizeof(oct_node));
    for (int k = 0; k < 8; k++) t->n[t->len].child[k] = -1;
    return (int32_t)t->len++;
}

static double luminance3(const double c[3]) { return 0.2126 * c[0] + 0.7152 * c[1] + 0.0722 * c[2]; }

static int32_t oct_build(oct_tree* t, const float* pts, const float* E, int64_t* idx, int64_t m, double area,
                         int64_t* scratch) {
    if (m == 0) return -1;
    double lo[3] = {INFTY, INFTY, INFTY}, hi[3] = {-INFTY, -INFTY, -INFTY};
    for (int64_t k = 0; k < m; k++)
        for (int d = 0; d < 3; d++) {
            double v = pts[3 * idx[k] + d];
            lo[d] = fmin(lo[d], v);
            hi[d] = fmax(hi[d], v);
        }
    int32_t id = oct_new(t);
    oct_node* nd = &t->n[id];
    for (int d = 0; d < 3; d++) {
        nd->lo[d] = lo[d];
        nd->hi[d] = hi[d];
    }
    if (m == 1 || (lo[0] == hi[0] && lo[1] == hi[1] && lo[2] == hi[2])) {
        /* one point, or points at one position (the reference would split
         * them forever): a leaf of all of them */
        nd->leaf = 1;
        for (int d = 0; d < 3; d++) nd->pos[d] = lo[d];
        for (int64_t k = 0; k < m; k++)
            for (int c = 0; c < 3; c++) nd->irad[c] += E[3 * idx[k] + c];
        for (int c = 0; c < 3; c++) nd->irad[c] /= (double)m;
        nd->area = area * (double)m;
        return id;
    }
    double mid[3];
    for (int d = 0; d < 3; d++) mid[d] = (lo[d] + hi[d]) * 0.5;
    int64_t cnt[8] = {0}, start[8];
    for (int64_t k = 0; k < m; k++) {
        const float* v = &pts[3 * idx[k]];
        int c = (v[0] < mid[0] ? 0 : 4) + (v[1] < mid[1] ? 0 : 2) + (v[2] < mid[2] ? 0 : 1);
        cnt[c]++;
    }
    start[0] = 0;
    for (int c = 1; c < 8; c++) start[c] = start[c - 1] + cnt[c - 1];
    int64_t fill[8];
    memcpy(fill, start, sizeof fill);
    for (int64_t k = 0; k < m; k++) {
        const float* v = &pts[3 * idx[k]];
        int c = (v[0] < mid[0] ? 0 : 4) + (v[1] < mid[1] ? 0 : 2) + (v[2] < mid[2] ? 0 : 1);
        scratch[fill[c]++] = idx[k];
    }
    memcpy(idx, scratch, sizeof(int64_t) * (size_t)m);
    int32_t ch[8];
    for (int c = 0; c < 8; c++) ch[c] = oct_build(t, pts, E, idx + start[c], cnt[c], area, scratch);
    nd = &t->n[id]; /* the array may have moved */
    double w = 0.0;
    int n_child = 0;
    for (int c = 0; c < 8; c++) {
        nd->child[c] = ch[c];
        if (ch[c] < 0) continue;
        const oct_node* cn = &t->n[ch[c]];
        double lw = luminance3(cn->irad);
        for (int d = 0; d < 3; d++) {
            nd->pos[d] += lw * cn->pos[d];
            nd->irad[d] += cn->irad[d];
        }
        nd->area += cn->area;
        w += lw;
        n_child++;
    }
    if (w > 0.0)
        for (int d = 0; d < 3; d++) nd->pos[d] /= w;
    for (int c = 0; c < 3; c++) nd->irad[c] /= (double)n_child;
    return id;
}

/* subsurface_integrator.cc:147-167 */
static void oct_query(const oct_tree* t, int32_t id, const double x[3], const double sa[3], const double ss[3],
                      double eta, double max_error, int32_t fast, double acc[3], int64_t* visited) {
    if (id < 0) return;
    const oct_node* nd = &t->n[id];
    (*visited)++;
    double dx = nd->pos[0] - x[0], dy = nd->pos[1] - x[1], dz = nd->pos[2] - x[2];
    double d2 = dx * dx + dy * dy + dz * dz;
    double dw = nd->area / d2;
    int inside = nd->lo[0] <= x[0] && x[0] <= nd->hi[0] && nd->lo[1] <= x[1] && x[1] <= nd->hi[1] &&
                 nd->lo[2] <= x[2] && x[2] <= nd->hi[2];
    if (nd->leaf || (dw < max_error && !inside)) {
        double rd[3];
        orc_dipole(sa, ss, eta, d2, fast, rd);
        for (int c = 0; c < 3; c++) acc[c] += rd[c] * nd->irad[c] * nd->area;
        return;
    }
    for (int c = 0; c < 8; c++) oct_query(t, nd->child[c], x, sa, ss, eta, max_error, fast, acc, visited);
}

void orc_dipole_octree(const double sigma_a[3], const double sigmap_s[3], double eta, const float* pts,
                       const float* E, int64_t n_pts, double area, const float* q, int64_t n_q, double max_error,
                       int32_t fast, double* out, int64_t* visited) {
    oct_tree t = {NULL, 0, 0};
    int64_t* idx = (int64_t*)malloc(sizeof(int64_t) * (size_t)(n_pts > 0 ? n_pts : 1));
    int64_t* scratch = (int64_t*)malloc(sizeof(int64_t) * (size_t)(n_pts > 0 ? n_pts : 1));
    for (int64_t k = 0; k < n_pts; k++) idx[k] = k;
    int32_t root = oct_build(&t, pts, E, idx, n_pts, area, scratch);
    double f = (1.0 - orc_dipole_fdr(eta)) / PI;
#pragma omp parallel for schedule(dynamic, 16)
    for (int64_t i = 0; i < n_q; i++) {
        double x[3] = {q[3 * i], q[3 * i + 1], q[3 * i + 2]};
        double acc[3] = {0.0, 0.0, 0.0};
        int64_t v = 0;
        oct_query(&t, root, x, sigma_a, sigmap_s, eta, max_error, fast, acc, &v);
        for (int c = 0; c < 3; c++) out[3 * i + c] = f * acc[c];
        if (visited) visited[i] = v;
    }
    free(t.n);
    free(idx);
    free(scratch);
}
