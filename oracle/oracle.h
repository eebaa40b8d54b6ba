/*
 * oracle.h -- CPU restatement of the tatsy-pppm probabilistic PPM pass.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP
 * implementation and the CPU baseline that bench.py times.  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it; the
 * product path never links, loads or calls it.
 *
 * Arithmetic is FP64 (as in the reference's Vector3D, sources/vector3d.h),
 * inputs are FP32 geometry widened exactly to FP64, and random numbers come
 * from the same counter-based generator as the GPU path, so both paths draw
 * identical uniforms for the same (seed, pass, stream, index, dimension).
 *
 * Each function cites the reference file:line it restates.  Parity status of
 * every function is listed in the header of oracle.c.
 */
#ifndef PPPM_ORACLE_H
#define PPPM_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Material flags: the reference BsdfType bits (sources/bsdf.h:19-26) plus an
 * emitter bit for the area light (the reference has no emissive material). */
#define ORC_LAMBERTIAN 0x01
#define ORC_SPECULAR   0x02
#define ORC_PHONG      0x04
#define ORC_REFRACTION 0x08
#define ORC_BSSRDF     0x10
#define ORC_EMITTER    0x40

/* Layout-identical to pppm_bsdf in include/pppm.h (64 bytes). */
typedef struct {
    int32_t type;
    float reflectance[3];
    float emission[3];
    float ior;
    float sigma_a[3];
    float sigmap_s[3];
    float eta;
    float pad;
} orc_bsdf;

/* Layout-identical to pppm_light: parallelogram corner + two edges. */
typedef struct {
    float corner[3];
    float edge_u[3];
    float edge_v[3];
    float emission[3];
} orc_light;

/* Layout-identical to pppm_camera. */
typedef struct {
    double center[3];
    double direction[3];
    double up[3];
    double fov_y_deg;
    int32_t width;
    int32_t height;
} orc_camera;

typedef struct orc_scene orc_scene;

/* ---- random numbers (Philox-4x32-10, Salmon et al. SC'11) ---- */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
/* n uniforms of the stream (seed, stream kind, pass, index, scene), dims 0..n-1 */
void orc_uniforms(uint64_t seed, uint32_t kind, uint32_t pass, uint32_t index,
                  uint32_t scene, int32_t n, double* out);

/* ---- scene + QBVH (sources/scene.cc, sources/qbvh_accel.cc) ---- */
orc_scene* orc_scene_create(const float* tri_xyz, const int32_t* bsdf_id, int64_t n,
                            const orc_bsdf* table, int32_t nb,
                            const orc_light* lights, int32_t nl);
void orc_scene_destroy(orc_scene* s);
/* rays: [n][8] = ox oy oz tmin dx dy dz tmax (FP32, as pppm_ray).  brute != 0
 * intersects every triangle (the reference test's ground truth). */
void orc_intersect_batch(const orc_scene* s, const float* rays, int64_t n, int32_t brute,
                         float* t_out, int32_t* tri_out);
int32_t orc_qbvh_stats(const orc_scene* s, int64_t* nodes, int64_t* leaves, int32_t* depth);

/* Environment map (sources/envmap.cc): rgb [h][w][3] lat-long, row 0 at +y;
 * w, h >= 64.  The importance map is the reference's (64 x 64 equal-area
 * cells, squared luminance of the mean lookup).  Returns 0, or -1 for a map
 * smaller than 64 x 64. */
int32_t orc_scene_set_envmap(orc_scene* s, const float* rgb, int32_t w, int32_t h);
/* The importance cells: pdf [4096] and their running sums cdf [4096] */
void orc_envmap_cells(const orc_scene* s, double* pdf, double* cdf);
/* Envmap::sampleFromDir of n directions [n][3] */
void orc_envmap_lookup(const orc_scene* s, const double* dirs, int64_t n, double* rgb);
/* The importance-map direction of uniforms r [n][3] (cell pick, u jitter,
 * v jitter; Envmap::samplePhoton's draws), L(w) / pdf(cell) and the cell. */
void orc_envmap_sample_dir(const orc_scene* s, const double* r, int64_t n, double* dir, double* l_over_pdf,
                           int32_t* cell);

/* ---- HashGrid<int> semantics (sources/hash_grid_detail.h) ---- */
uint32_t orc_hash(int32_t ix, int32_t iy, int32_t iz, int32_t hash_size);
/* Point insertion: add(i, p_i, p_i) for i = 0..n-1, then the concatenation
 * of buckets 0..S-1 into sorted_idx with [cell_start, cell_end) ranges. */
int32_t orc_hashgrid_points(const float* pos, int64_t n, int32_t hash_size, double hash_scale,
                            const double bbox_min[3], int32_t* sorted_idx,
                            int32_t* cell_start, int32_t* cell_end);
/* Box insertion add(i, bmin_i, bmax_i); returns the number of (item, cell)
 * pairs, or -1 if it exceeds cap. */
int64_t orc_hashgrid_boxes(const double* box_min, const double* box_max, int64_t n,
                           int32_t hash_size, double hash_scale, const double bbox_min[3],
                           int32_t* sorted_items, int64_t cap,
                           int32_t* cell_start, int32_t* cell_end);

/* ---- radius schedule (sources/ppm_probabilistic.cc:94) ---- */
/* literal=0: Knaus-Zwicker r2_{i+1} = r2_i (i+alpha)/(i+1);
 * literal=1: the reference's r_{t+1} = (t+1)/(t+alpha) r_t.  out[i] = r_{i+1}. */
void orc_radius_schedule(double r1, double alpha, int32_t n, int32_t literal, double* out);

/* ---- PPM pass pieces (sources/ppm_probabilistic.cc) ---- */
typedef struct {
    int64_t photons;        /* emitted photons this pass */
    int32_t pass_index;     /* 1-based pass number i */
    int32_t max_bounces;    /* reference bounceLimit = 64 */
    double radius;          /* r_i */
    double ray_eps;         /* spawn offset along the surface normal */
    uint64_t seed;
    uint32_t scene_id;
    int32_t pad;
} orc_pass_params;

/* Photon pass. Deposits in (photon index, bounce) order.  pos/flux/normal are
 * [cap][3], key (may be NULL) = photon index << 6 | bounce.  Returns the
 * deposit count (may exceed cap; then only cap are written). */
int64_t orc_trace_photons(const orc_scene* s, const orc_pass_params* p,
                          float* pos, float* flux, float* normal, uint32_t* key, int64_t cap);

/* Learned emission proposal (pppm_device.h, include/pppm.h): per-anchor
 * inclusive CDF rows [side^2][1024] over 32 x 32 equal-area direction cells,
 * and the weight lambda of the cosine component. */
typedef struct {
    const float* cdf;
    double lambda;
    int32_t side;
    int32_t pad;
} orc_proposal;
/* Photon pass emitting from the mixture (pr NULL: plain cosine emission). */
int64_t orc_trace_photons_prop(const orc_scene* s, const orc_pass_params* p, const orc_proposal* pr,
                               float* pos, float* flux, float* normal, uint32_t* key, int64_t cap);
/* Emission of photon idx of a pass about light normal nl: light-sample
 * coordinates, direction, flux factor (cos / pi) / p(w) and direction cell. */
void orc_emission(const orc_proposal* pr, const orc_pass_params* p, int64_t idx, const double nl[3],
                  double out_uv[2], double out_dir[3], double* weight, uint32_t* cell);

/* Dipole BSSRDF (sources/bssrdf.cc:33-104): F_dr, R_d(d2) per channel
 * (clamped at 0; fast = 1: the reference's Schraudolph exp), and the surface
 * sum (1 - F_dr) / pi * area * sum_j R_d(|q_i - p_j|^2) E_j in FP64. */
double orc_dipole_fdr(double eta);
void orc_dipole(const double sigma_a[3], const double sigmap_s[3], double eta, double d2, int32_t fast,
                double out[3]);
void orc_dipole_sum(const double sigma_a[3], const double sigmap_s[3], double eta, const float* pts,
                    const float* E, int64_t n_pts, double area, const float* q, int64_t n_q, double* out);

/* The same sum over an octree cut (sources/subsurface_integrator.cc:69-167):
 * a node stands for its points when it is a leaf, or when area / d^2 <
 * max_error and q lies outside its box; node values follow the reference
 * (position: luminance-weighted mean of the children's, area: their sum,
 * irradiance: the mean of the children's).  Points at one position share a
 * leaf.  fast = 1: the reference's exponential in R_d.  visited (may be
 * null): nodes visited per query. */
void orc_dipole_octree(const double sigma_a[3], const double sigmap_s[3], double eta, const float* pts,
                       const float* E, int64_t n_pts, double area, const float* q, int64_t n_q, double max_error,
                       int32_t fast, double* out, int64_t* visited);

/* Eye pass: one path per pixel.  For each pixel writes either a hit point
 * (has_hit=1: pos, oriented normal, weight = throughput * rho / pi) or a
 * direct contribution (emission seen through specular chains). */
void orc_eye_pass(const orc_scene* s, const orc_camera* cam, const orc_pass_params* p,
                  int32_t* has_hit, float* hit_pos, float* hit_nrm, float* hit_w,
                  double* direct_rgb);

/* Fixed-radius cone-filtered gather (P-KZ form of ppm_probabilistic.cc:300-334)
 * of n_hit hit points over n_ph photons through the hash grid. out_rgb [n_hit][3].
 * counters (may be NULL): [0] buckets visited, [1] candidates, [2] in radius. */
void orc_gather(const float* ph_pos, const float* ph_flux, const float* ph_nrm, int64_t n_ph,
                const float* hit_pos, const float* hit_nrm, const float* hit_w, int64_t n_hit,
                double radius, int32_t hash_size, const double bbox_min[3],
                double* out_rgb, int64_t* counters);

/* Full P-KZ render: passes first_pass..first_pass+n_passes-1, radius r1
 * shrinking by alpha, FP64 image (top row first) averaged over passes.
 * stats (may be NULL): [0] total deposits, [1] total hit points. */
int32_t orc_render(const orc_scene* s, const orc_camera* cam, int64_t photons_per_pass,
                   int32_t n_passes, int32_t first_pass, double r1, double alpha,
                   double ray_eps, uint64_t seed, uint32_t scene_id, int32_t hash_size,
                   double* image, int64_t* stats);

/* Options of orc_render_ex (NULL: plain emission, no subsurface). */
typedef struct {
    const orc_proposal* prop;  /* learned-proposal emission (NULL: cosine emission) */
    const float* sss_pts;      /* [n][3] surface points of the BSSRDF materials (NULL: no subsurface) */
    const float* sss_nrm;      /* [n][3] their outward normals */
    int64_t sss_n;
    double sss_area;           /* surface area each point stands for */
} orc_render_opts;

/* orc_render with a proposal and/or the subsurface estimate: every pass also
 * stores the photons reaching BSSRDF surfaces from outside, gathers their
 * irradiance E_j at the points (weight 1 / pi), and an eye path transmitted
 * into a BSSRDF body adds W (1 - F_dr) / pi * area * sum_j R_d(|x - p_j|^2) E_j.
 * stats (may be NULL): [0] deposits, [1] hit points, [2] irradiance photons,
 * [3] transmitted eye paths. */
int32_t orc_render_ex(const orc_scene* s, const orc_camera* cam, int64_t photons_per_pass,
                      int32_t n_passes, int32_t first_pass, double r1, double alpha,
                      double ray_eps, uint64_t seed, uint32_t scene_id, int32_t hash_size,
                      const orc_render_opts* opts, double* image, int64_t* stats);

/* Irradiance photons of a pass (BSSRDF surfaces, arriving from outside), in
 * (photon index, bounce) order; returns the count (only cap written). */
int64_t orc_trace_irradiance_photons(const orc_scene* s, const orc_pass_params* p, float* pos, float* flux,
                                     float* normal, uint32_t* key, int64_t cap);

/* Progressive photon mapping with per-point radii (sources/
 * progressive_photon_mapping.cc:22-287): render points per pixel, updated
 * photon by photon in a fixed order.  stats (may be NULL): [0] accepted
 * photon updates. */
int32_t orc_render_hachisuka(const orc_scene* s, const orc_camera* cam, int64_t photons, int32_t n_passes,
                             double alpha, double ray_eps, uint64_t seed, uint32_t scene_id, int32_t hash_size,
                             double* image, int64_t* stats);

/* Independent FP64 path tracer with next-event estimation to the area
 * lights: image [height][width][3] (top row first) = mean over samples
 * first_sample .. first_sample + spp - 1 of each pixel's radiance. */
void orc_pathtrace(const orc_scene* s, const orc_camera* cam, int32_t spp, int32_t first_sample,
                   int32_t max_bounces, double ray_eps, uint64_t seed, uint32_t scene_id, double* image);

/* Single-sample pieces, exported to pin them against the compiled reference:
 * sampler.cc:17-32, reflectance.h:7-34, brdf.cc, perspective_camera.cc:42-46,
 * triangle.cc:53-78. */
void orc_on_hemisphere(const double n[3], double r1, double r2, double out[3]);
int32_t orc_total_reflection(int32_t into, const double in[3], const double n[3], const double on[3],
                             double refl[3], double refr[3], double* fre, double* ftr);
void orc_bsdf_sample(int32_t type, const double rho[3], const double in[3], const double n[3],
                     double r1, double r2, double out[3], double* pdf);
void orc_camera_ray(const orc_camera* c, double px, double py_up, double o[3], double d[3]);
int32_t orc_tri_intersect(const double tri[9], const double o[3], const double d[3], double* t);

/* Threads the OpenMP runtime will use. */
int32_t orc_num_threads(void);

#ifdef __cplusplus
}
#endif
#endif
