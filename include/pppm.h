/*
 * pppm.h -- C ABI of the MI355X progressive photon mapping library
 * (libpppm_hip.so, built for gfx950).
 *
 * Every entry point returns an int status (PPPM_OK = 0); the message of the
 * last failure on a context is pppm_last_error(ctx).  Nothing aborts across
 * the ABI.  Host buffers are owned by the caller; device memory is owned by
 * the context.  One context per host thread; calls are synchronous.
 *
 * Each declaration names the reference interface it replaces
 * (tatsy/tatsy-pppm, paths relative to the reference root).
 */
#ifndef PPPM_H
#define PPPM_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PPPM_OK 0
#define PPPM_ERR_INVALID 1  /* bad argument */
#define PPPM_ERR_HIP 2      /* HIP runtime error (no device, launch failure ...) */
#define PPPM_ERR_STATE 3    /* call out of order (no scene, no BVH, no camera ...) */
#define PPPM_ERR_CAPACITY 4 /* photon buffer too small; stats report the size needed */
#define PPPM_ERR_NOMEM 5

/* Material flags: sources/bsdf.h:19-26 BsdfType, plus an emitter bit. */
#define PPPM_BSDF_LAMBERTIAN 0x01
#define PPPM_BSDF_SPECULAR 0x02
#define PPPM_BSDF_PHONG 0x04 /* not supported: rejected by pppm_scene_upload */
#define PPPM_BSDF_REFRACTION 0x08
#define PPPM_BSDF_BSSRDF 0x10
#define PPPM_BSDF_EMITTER 0x40

/* One material (64 bytes).  Replaces the polymorphic BSDF handle of
 * sources/bsdf.h:32-59 (reflectance() + type()) and the dipole parameters of
 * sources/bssrdf.h:47-66. */
typedef struct {
    int32_t type;
    float reflectance[3];
    float emission[3]; /* radiance of an emitter's front face */
    float ior;         /* refraction: 1.5 in the reference (sources/common.h:16) */
    float sigma_a[3];
    float sigmap_s[3];
    float eta;
    float pad;
} pppm_bsdf;

/* Parallelogram area light: corner + two edges; emits on the side of
 * edge_u x edge_v.  (The reference's only emitter is Envmap::samplePhoton,
 * sources/envmap.cc:62-86; this is the emission contract the configs need.) */
typedef struct {
    float corner[3];
    float edge_u[3];
    float edge_v[3];
    float emission[3];
} pppm_light;

/* Pinhole camera (sources/perspective_camera.h:24-31): vertical field of view
 * in degrees, orthonormal image basis, image row 0 at the top. */
typedef struct {
    double center[3];
    double direction[3];
    double up[3];
    double fov_y_deg;
    int32_t width;
    int32_t height;
} pppm_camera;

/* Ray for batch queries (sources/ray.h:16-36) and its hit (ray.h:38-79):
 * tri = triangle index in upload order, -1 on a miss. */
typedef struct {
    float origin[3];
    float tmin;
    float dir[3];
    float tmax;
} pppm_ray;

typedef struct {
    float t;
    int32_t tri;
    float u, v;
} pppm_hit;

typedef struct {
    int64_t nodes;     /* 4-wide nodes */
    int64_t leaves;
    int32_t depth;     /* root = 1 */
    int32_t max_leaf;  /* triangles per leaf limit */
    double build_ms;
    int32_t builder;   /* PPPM_BVH_LBVH or PPPM_BVH_PLOC: the binary tree that was collapsed */
    int32_t reserved;
} pppm_bvh_info;

/* Per-render settings (RenderParameters, sources/render_parameters.h:4-44,
 * plus the knobs the reference hard-codes: ALPHA, ppm_probabilistic.cc:9;
 * initial radius, :55). */
typedef struct {
    int64_t photons;       /* emitted photons per pass (RenderParameters::photons) */
    int32_t passes;        /* RenderParameters::spp: number of passes */
    int32_t max_bounces;   /* bounceLimit, 64 in the reference */
    double alpha;          /* 0.7 */
    double radius1;        /* r_1, scene units */
    double ray_eps;        /* continuation-ray offset along the normal */
    uint64_t seed;
    uint32_t scene_id;     /* RNG stream id of this scene */
    int32_t hash_size;     /* 0: the number of stored photons of each pass */
    double capacity_factor;/* photon buffer = photons * factor records (0: 8) */
    int32_t gather_mode;   /* PPPM_GATHER_AUTO (0, the default: resolves to LANES), LANES, TILED, WAVE or BLOCKS */
    int32_t method;        /* PPPM_METHOD_PROBABILISTIC (0) or PPPM_METHOD_HACHISUKA */
} pppm_render_params;

/* Render methods.  PROBABILISTIC: ProgressivePhotonMappingProb
 * (sources/ppm_probabilistic.cc), one global radius shrinking per pass (the
 * P-KZ form).  HACHISUKA: ProgressivePhotonMapping
 * (sources/progressive_photon_mapping.cc:22-287), a render point per pixel
 * with its own radius, photon count and flux, updated photon by photon with
 * g = (n alpha + alpha) / (n alpha + 1); radius1 is unused (each pass's initial
 * radius comes from the hit points' extent, :116-125).  No learned proposal
 * and no BSSRDF materials with HACHISUKA; pppm_render_passes runs its passes
 * one by one. */
#define PPPM_METHOD_PROBABILISTIC 0
#define PPPM_METHOD_HACHISUKA 1

#define PPPM_GATHER_AUTO 0  /* the kernel measured fastest on gfx950 (currently LANES) */
#define PPPM_GATHER_LANES 1 /* one lane per Morton-sorted hit point over its own buckets */
#define PPPM_GATHER_TILED 2 /* LDS-tiled: a wave's union of buckets staged in 64-photon tiles */
#define PPPM_GATHER_WAVE 3  /* wave-uniform bucket walk, photons through scalar loads */
#define PPPM_GATHER_BLOCKS 4 /* wave-uniform walk of the 32-record blocks some lane's ball meets, staged in LDS */

/* Timings and counters of one pass. */
typedef struct {
    int32_t pass_index;
    int32_t hash_size;
    double radius;
    int64_t photons_emitted;
    int64_t photons_stored;
    int64_t photon_segments;   /* ray queries of the photon pass */
    int64_t eye_segments;      /* ray queries of the eye pass */
    int64_t hitpoints;
    int64_t gather_buckets;    /* sum over hit points of de-duplicated buckets */
    int64_t gather_candidates; /* sum of photon records in them (the reference's candidates) */
    int64_t gather_inradius;
    double photon_ms, order_ms, grid_ms, eye_ms, gather_ms, total_ms;
    double hitsort_ms;         /* Morton sort of the hit points (tiled gather) */
    /* learned emission proposal (zero when it is off) */
    double proposal_ms;        /* contribution records + training + tables */
    double train_loss;         /* mean weighted cross entropy of the last retrain */
    int64_t proposal_records;  /* records written this pass */
    int32_t proposal_active;   /* photons of this pass used the learned mixture */
    int32_t retrained;         /* the MLP was retrained after this pass */
    /* subsurface scattering (zero without BSSRDF materials) */
    double irradiance_ms;      /* irradiance photons: order, grid, gather at the surface points */
    double dipole_ms;          /* dipole diffusion sum of the transmitted eye paths */
    int64_t irradiance_photons;/* photons stored on BSSRDF surfaces */
    int64_t sss_hits;          /* eye paths ending in the dipole estimate */
    int64_t sss_points;        /* surface points of the irradiance cache */
    /* gather work beyond the reference convention (gather_candidates) */
    int64_t gather_tested;     /* photon records the kernel distance-tested */
    int64_t gather_blocks;     /* 32-record block boxes tested (lanes kernel) */
    /* traversal work (zero unless pppm_count_traversal is on) */
    int64_t photon_nodes;      /* inner BVH nodes visited by the photon pass */
    int64_t photon_tris;       /* triangles tested by the photon pass */
    int64_t eye_nodes, eye_tris;
    int64_t gather_staged;     /* records staged in LDS, once per wave (blocks kernel): x 64 lane tests */
} pppm_pass_stats;

/* ---- context ---- */
int pppm_create(void** ctx, int32_t device);
int pppm_destroy(void* ctx);
const char* pppm_last_error(const void* ctx);
int pppm_device_info(void* ctx, char* name, int32_t name_len, int32_t* cu_count,
                     int64_t* hbm_bytes);

/* ---- scene: replaces Scene::add / setAccelerator (sources/scene.h:38-56,
 * sources/scene.cc:35-38) and QBVHAccel::construct (sources/qbvh_accel.h:49).
 * nl may be 0 when an environment map provides the light. ---- */
int pppm_scene_upload(void* ctx, const float* tri_xyz /* [n][9] */, const int32_t* bsdf_id,
                      int64_t n, const pppm_bsdf* table, int32_t nb, const pppm_light* lights,
                      int32_t nl);
/* GPU BVH build: a binary tree over the Morton-sorted triangles collapsed to
 * 4-wide 128-byte nodes; pppm_bvh_build uses PPPM_BVH_DEFAULT. */
#define PPPM_BVH_DEFAULT 0 /* the builder whose trees traced fastest on gfx950 */
#define PPPM_BVH_LBVH 1    /* radix tree over the Morton codes (Karras 2012) */
#define PPPM_BVH_PLOC 2    /* locally-ordered clustering on surface area (Meister and Bittner 2018);
                            * falls back to LBVH when its tree would be deeper than the traversal
                            * stack allows or its clustering would run past 64 + 8 log2(n) rounds */
int pppm_bvh_build(void* ctx, pppm_bvh_info* info);
int pppm_bvh_build_with(void* ctx, int32_t builder, pppm_bvh_info* info);
/* Batch closest hit: replaces Scene::intersect (sources/scene.cc:40-47) and
 * QBVHAccel::intersect (sources/qbvh_accel.cc:205-285). */
int pppm_intersect(void* ctx, const pppm_ray* rays, int64_t n, pppm_hit* hits);

/* One 4-wide BVH node as the GPU builds it (128 bytes): the four child boxes
 * in SoA FP32 (an empty slot has lo > hi), then four child refs: ref >= 0 is
 * an inner node index, ref < 0 a leaf whose ~ref = (first << 3) | (count - 1)
 * indexes the leaf-ordered triangles, 0x7FFFFFFF an empty slot. */
typedef struct {
    float lox[4], hix[4], loy[4], hiy[4], loz[4], hiz[4];
    int32_t child[4];
    int32_t pad[4];
} pppm_bvh_node;

/* Host copy of the built BVH, for scalar re-entrant traversal on the CPU
 * (the reference's QBVHAccel::intersect is const and called concurrently,
 * sources/qbvh_accel.cc:205-285, ppm_probabilistic.cc:157, 256): nodes
 * [n_nodes], leaf-ordered triangles tris [n_tris][12] (v0 xyz + original
 * triangle index as float bits, v1 xyz + 0, v2 xyz + 0) and the root ref.
 * With null buffers only the counts are returned. */
int pppm_bvh_download(void* ctx, pppm_bvh_node* nodes, int64_t node_cap, float* tris, int64_t tri_cap,
                      int64_t* n_nodes, int64_t* n_tris, int32_t* root);

/* ---- hash grid: replaces HashGrid<Ty> (sources/hash_grid.h:8-39,
 * sources/hash_grid_detail.h:26-72) ---- */
/* Point insertion add(i, p_i, p_i), i = 0..n-1: sorted_idx[cell_start[h] ..
 * cell_end[h]) is bucket h in insertion order (empty bucket: start = end =
 * number of items in lower buckets). */
int pppm_hashgrid_build(void* ctx, const float* pos /* [n][3] */, int64_t n, int32_t hash_size,
                        double hash_scale, const double bbox_min[3], int32_t* sorted_idx,
                        int32_t* cell_start, int32_t* cell_end);
/* The point-insertion grid of pppm_hashgrid_build without any allocation.
 * No counterpart in the reference: for callers that cannot allocate device
 * memory (code under stream capture, tight serving loops, harnesses that
 * forbid hipMalloc).  pos, sorted_idx, cell_start and cell_end are DEVICE
 * buffers the caller owns (same layouts as pppm_hashgrid_build); bbox_min is
 * host memory.  Intermediates live in one fixed array per GPU (about 21 MB,
 * held until the process ends), so n <= PPPM_NOALLOC_MAX_POINTS and
 * hash_size <= PPPM_NOALLOC_MAX_BUCKETS (PPPM_ERR_CAPACITY otherwise).
 * Calls that overlap on one GPU race: every call on a GPU shares its array.
 * Synchronous on the context's stream. */
#define PPPM_NOALLOC_MAX_POINTS (1 << 20)
#define PPPM_NOALLOC_MAX_BUCKETS (1 << 20)
int pppm_hashgrid_build_noalloc(void* ctx, const float* pos, int64_t n, int32_t hash_size, double hash_scale,
                                const double bbox_min[3], int32_t* sorted_idx, int32_t* cell_start,
                                int32_t* cell_end);
/* Box insertion add(i, box_min_i, box_max_i): every (item, cell) pair in the
 * reference's z, y, x loop order.  *n_pairs returns the pair count; fails
 * with PPPM_ERR_CAPACITY if it exceeds cap. */
int pppm_hashgrid_build_boxes(void* ctx, const double* box_min /* [n][3] */,
                              const double* box_max, int64_t n, int32_t hash_size,
                              double hash_scale, const double bbox_min[3], int32_t* sorted_items,
                              int64_t cap, int64_t* n_pairs, int32_t* cell_start,
                              int32_t* cell_end);

/* ---- environment map: replaces Envmap (sources/envmap.cc:20-163) as an
 * emitter beside the area lights, and the radiance of eye rays that leave the
 * scene (ppm_probabilistic.cc:248-251) ---- */
/* rgb [height][width][3] linear radiance, lat-long: row 0 looks along +y,
 * column x at phi = 2 pi (x + 0.5) / width from +x towards +z; width,
 * height >= 64.  rgb = NULL removes the map.  Set after pppm_scene_upload
 * (which removes it) and before pppm_render_begin.  Photons take it by power
 * against the area lights: a direction from the reference's 64 x 64
 * importance cells, a point on a disc of the scene's bounding-sphere radius
 * facing that direction, flux L pi R^2 / (p(w) N P) with the exact
 * solid-angle density p(w) of the cell (the reference's flux scale omits the
 * disc area, envmap.cc:81-83). */
int pppm_envmap_set(void* ctx, const float* rgb, int32_t width, int32_t height);
/* The importance cells alone, on the host (no device): pdf [4096] and running
 * sums cdf [4096] of Envmap::createImportanceMap (envmap.cc:88-163). */
int pppm_envmap_importance(const float* rgb, int32_t width, int32_t height, double* pdf, double* cdf);

/* ---- radius schedule (sources/ppm_probabilistic.cc:94), FP64 on the host ---- */
/* literal = 0: Knaus-Zwicker r2_{i+1} = r2_i (i+alpha)/(i+1); literal = 1: the
 * reference's r_{t+1} = (t+1)/(t+alpha) r_t.  out[k] = r_{k+1}. */
int pppm_radius_schedule(double r1, double alpha, int32_t n, int32_t literal, double* out);

/* ---- renderer: replaces ProgressivePhotonMappingProb::render
 * (sources/ppm_probabilistic.h:35, ppm_probabilistic.cc:24-121) split into
 * begin / pass / read so the caller owns the pass loop ---- */
int pppm_camera_set(void* ctx, const pppm_camera* cam);
int pppm_render_begin(void* ctx, const pppm_render_params* params);
/* Runs pass i = (passes done) + 1 with r_i from the schedule. */
int pppm_render_pass(void* ctx, pppm_pass_stats* stats);
/* Runs `count` passes pipelined on three streams: the gather of pass i
 * overlaps the photon tracing of pass i + 1 (the image equals `count`
 * pppm_render_pass calls bit for bit).  stats (may be NULL) gets one entry per
 * pass.  With the learned proposal or BSSRDF materials the passes run one by
 * one. */
int pppm_render_passes(void* ctx, int32_t count, pppm_pass_stats* stats);
/* Measurement hook, no reference counterpart: with enable = 1 the following
 * passes run kernel variants that also count BVH nodes visited and triangles
 * tested per pass (stats photon_nodes ... eye_tris); the images are the same.
 * For measuring traversal bytes; the counting costs a little time. */
int pppm_count_traversal(void* ctx, int32_t enable);
/* FP64 image [height][width][3], row 0 at the top, averaged over the passes
 * done: (sum - compensation) / passes. */
int pppm_image_read(void* ctx, double* rgb, int64_t n_values);
/* The current image (as pppm_image_read averages it) encoded on the device
 * and written to path by its extension: .png / .bmp 8-bit RGB after
 * c^(1 / gamma) (gamma <= 0: none) clamped to [0, 1] and truncated to a byte
 * (Image::gamma + Image::savePng / saveBmp, sources/image.cc:202-213,
 * 484-496, 530-533); .hdr Radiance RGBE (Image::saveHdr, :436-482); .pfm
 * linear FP32.  Only the encoded bytes cross to the host. */
int pppm_image_save(void* ctx, const char* path, double gamma);
/* Host image files without a device: rgb [height][width][3] FP64, top row
 * first, written by extension as pppm_image_save writes them. */
int pppm_image_write(const char* path, const double* rgb, int32_t width, int32_t height, double gamma);
/* Reads .hdr (Radiance RGBE, run-length or flat scanlines; Image::loadHdr,
 * sources/image.cc:296-434), .pfm or 24-bit .bmp into an FP64 [height][width][3]
 * buffer the library allocates (release with pppm_image_free). */
int pppm_image_load(const char* path, double** rgb, int32_t* width, int32_t* height, char* err,
                    int32_t err_len);
void pppm_image_free(double* rgb);
/* Test hooks for the last pass: photons in (photon index, bounce) order,
 * key = photon index << 6 | bounce, and the hash grid over them. */
int pppm_last_photons(void* ctx, float* pos, float* flux, float* normal, uint32_t* key,
                      int64_t cap, int64_t* n);
int pppm_last_grid(void* ctx, int32_t* sorted_idx, int32_t* cell_start, int32_t* cell_end,
                   int32_t hash_size, double bbox_min[3], double* hash_scale);
/* Hit points of the last pass: pos, oriented normal, weight, pixel id. */
int pppm_last_hitpoints(void* ctx, float* pos, float* normal, float* weight, int32_t* pixel,
                        int64_t cap, int64_t* n);
/* Gather only: given photons and hit points, the per-hit-point radiance of
 * the fixed-radius cone-filtered estimate (sources/ppm_probabilistic.cc:300-334). */
int pppm_gather(void* ctx, const float* ph_pos, const float* ph_flux, const float* ph_nrm,
                int64_t n_ph, const float* hit_pos, const float* hit_nrm, const float* hit_w,
                int64_t n_hit, double radius, int32_t hash_size, const double bbox_min[3],
                int32_t gather_mode, float* out_rgb, pppm_pass_stats* stats);

/* ---- measurement ---- */
/* Plain device-to-device copy of `bytes` (16-byte loads and stores, no
 * reference counterpart): the known-traffic kernel that scales profiler
 * FETCH_SIZE / WRITE_SIZE counters, and a measured copy bandwidth. */
int pppm_stream_copy(void* ctx, int64_t bytes, double* ms);
/* Accumulator test hook (no reference counterpart): lane k (< lanes <= 64)
 * adds values[(i + k) % n], i = 0 .. n-1, with the image's FP32 compensated
 * (Kahan) addition; sum[k] and comp[k] receive the pair, whose value is
 * sum - comp (as pppm_image_read reads it). */
int pppm_kahan_probe(void* ctx, const float* values, int64_t n, int32_t lanes, float* sum, float* comp);
/* Issue-rate calibration (no reference counterpart): cycles per wave64
 * instruction on one SIMD at the reported clock for a full-rate v_fma_f32
 * and for v_exp_f32, measured with independent chains; the roofline of the
 * dipole sum is built from them. */
int pppm_issue_probe(void* ctx, double* valu_cycles, double* trans_cycles, double* clock_hz);

/* ---- subsurface scattering: replaces SubsurfaceIntegrator (sources/
 * subsurface_integrator.cc:180-365, called per pass from ppm_probabilistic.cc:
 * 83-85 and 272-296) with DipoleBSSRDF (sources/bssrdf.cc:33-104).  A BSSRDF
 * material (PPPM_BSDF_REFRACTION | PPPM_BSDF_BSSRDF, sigma_a, sigmap_s, eta)
 * gets fixed surface points (each standing for an equal share of the area);
 * every pass stores the photons that reach those surfaces, estimates the
 * irradiance E_j at the points with the fixed-radius gather, and adds
 * (1 - F_dr) / pi * sum_j R_d(|x - x_j|^2) E_j A_j for eye paths transmitted
 * into the surface.  The exponential is the true one, not the reference's
 * fast_exp (sources/vector3d.cc:11-16). ---- */
/* Surface points of later renders (0: the reference's density, at most 2^20);
 * takes effect at pppm_render_begin. */
int pppm_subsurface_configure(void* ctx, int64_t points);
/* The sum over the surface points: max_error 0 (default) sums every point;
 * max_error > 0 walks the reference's octree cut, taking a node whole when
 * its area / d^2 < max_error and x lies outside its box
 * (sources/subsurface_integrator.cc:69-167; the reference passes 0.05 at
 * ppm_probabilistic.cc:84).  Takes effect at
 * pppm_render_begin, and for pppm_dipole_eval at once. */
int pppm_subsurface_cut(void* ctx, double max_error);
/* Points of the current render [n][3], their normals, and the area per point. */
int pppm_subsurface_points(void* ctx, float* pos, float* normal, int64_t cap, int64_t* n, double* area);
/* Irradiance E_j [n][3] of the last pass. */
int pppm_subsurface_irradiance(void* ctx, float* E, int64_t cap, int64_t* n);
/* Photons stored on BSSRDF surfaces in the last pass (append order). */
int pppm_last_irradiance_photons(void* ctx, float* pos, float* flux, float* normal, uint32_t* key, int64_t cap,
                                 int64_t* n);
/* The dipole sum alone (DipoleBSSRDF::operator() of sources/bssrdf.cc:92-104
 * summed as in subsurface_integrator.cc:147-167): out [m][3] =
 * (1 - F_dr) / pi * area * sum_j R_d(|q - p_j|^2) E_j for the material's
 * sigma_a, sigmap_s and eta. */
int pppm_dipole_eval(void* ctx, const pppm_bsdf* material, const float* pts, const float* E, int64_t n_pts,
                     double area, const float* query, int64_t n_query, float* out);

/* ---- learned emission proposal.  No reference counterpart: it replaces the
 * cosine emission of sources/sampler.cc:17-32 (the light contract of
 * sources/envmap.cc:62-86, used by sources/ppm_probabilistic.cc:136-146) by
 * the defensive mixture  p(w | x) = lambda cos(theta) / pi
 * + (1 - lambda) q_a(x)[k(w)] / (2 pi / 1024)  over a 32 x 32 equal-area grid
 * of directions about the light normal (cell k = kz * 32 + kphi with
 * 1 - cos(theta) in [kz / 32, (kz + 1) / 32) and phi in [kphi, kphi + 1) 2 pi / 32),
 * where a(x) is the anchor cell of the light sample x on an anchors x anchors
 * grid.  Photon flux carries the exact ratio (cos(theta) / pi) / p(w), so the
 * estimator stays unbiased; lambda = 1 is plain emission bit for bit.
 * q comes from an MLP (Fourier encoding of the light-sample position, 32 ->
 * 64 -> 64 -> 64 -> 1024, ReLU, BF16 MFMA with FP32 accumulation, Adam in
 * FP32) retrained every train_interval passes on contribution records of the
 * last passes.  One light only. ---- */
#define PPPM_PROPOSAL_CELLS 1024
#define PPPM_MLP_PARAMS 76992  /* W1 [64][32] b1 W2 [64][64] b2 W3 [64][64] b3 W4 [1024][64] b4 */
typedef struct {
    double lambda;            /* weight of the cosine component, (0, 1] */
    double learning_rate;     /* Adam step size */
    int64_t ring_capacity;    /* contribution records kept (the most recent) */
    int32_t anchors;          /* anchors per light edge (1..32) */
    int32_t train_interval;   /* retrain after every this many passes */
    int32_t train_steps;      /* Adam steps per retrain */
    int32_t batch;            /* records per step, a multiple of 256 (<= 65536) */
    uint64_t seed;            /* weight initialisation and mini-batch draws */
} pppm_proposal_params;
/* After pppm_render_begin: starts recording contributions and training; the
 * first retrain happens after pass train_interval, and passes after it sample
 * from the mixture.  pppm_render_begin switches the proposal off. */
int pppm_proposal_enable(void* ctx, const pppm_proposal_params* params);
/* Fixed tables instead of the MLP (tests, or a proposal trained elsewhere):
 * prob [anchors^2][1024] cell probabilities per anchor (row-major by the
 * anchor's (edge_u, edge_v) index), normalised on the device. */
int pppm_proposal_set_table(void* ctx, const float* prob, int32_t anchors, double lambda);
/* Current per-anchor inclusive CDFs [anchors^2][1024]. */
int pppm_proposal_get_table(void* ctx, float* cdf, int64_t n_values, int32_t* anchors);
/* The contribution records held (oldest first): [n][4] = u, v, cell, c. */
int pppm_proposal_records(void* ctx, float* rec, int64_t cap, int64_t* n);
/* MLP test hooks: FP32 master parameters; softmax probabilities [n][1024] at
 * light-sample coordinates uv [n][2]; one step on given records [n][4]
 * (n a multiple of 256): the reduced gradient (grad may be NULL), the weighted
 * cross entropy, and an Adam step when learning_rate > 0 (step = Adam's t). */
int pppm_mlp_set_params(void* ctx, const float* params, int64_t n);
int pppm_mlp_get_params(void* ctx, float* params, int64_t n);
int pppm_mlp_forward(void* ctx, const float* uv, int64_t n, float* prob);
int pppm_mlp_train_step(void* ctx, const float* rec, int64_t n, double learning_rate, int32_t step,
                        float* grad, double* loss);

/* ---- batched scenes: many renders in one hipGraph-captured pass loop.  No
 * reference counterpart (the reference renders one scene per process,
 * sources/main/appmain.cc:11-26); each scene's pass is pppm_render_pass's
 * computation (same image bit for bit, for the same hash size and gather
 * kernel), enqueued without host synchronisation: counts stay on the device,
 * sorts run over the buffer capacity, pass index and radius live in device
 * memory.  One pass of every scene is captured once into a hipGraph whose
 * branches (streams) run concurrently, scenes spread over them longest
 * first; each replay renders one more pass of every scene.  Plain PPM. ---- */
int pppm_batch_create(void** batch, int32_t device, int32_t branches);
int pppm_batch_destroy(void* batch);
const char* pppm_batch_last_error(const void* batch);
/* A scene with its camera and render parameters (hash_size > 0 required). */
int pppm_batch_add(void* batch, const float* tri_xyz, const int32_t* bsdf_id, int64_t n, const pppm_bsdf* table,
                   int32_t nb, const pppm_light* lights, int32_t nl, const pppm_camera* camera,
                   const pppm_render_params* params, int32_t* index);
/* Autotuner: per scene, times one pass of each gather kernel (tune_gather)
 * with each photon-kernel grid of 1, 2, 4 and 8 workgroups of 256 lanes per
 * CU (tune_grid), then, at the fastest of those, the photon traversal's lanes
 * kept in flight before a wave regenerates paths (8, 16, 24, 32, 48;
 * tune_grid); keeps the fastest.  None changes the image.  The renders
 * restart. */
int pppm_batch_tune(void* batch, int32_t tune_gather, int32_t tune_grid);
/* Runs `passes` more passes of every scene (graph replays); *ms = GPU time. */
int pppm_batch_run(void* batch, int32_t passes, double* ms);
int pppm_batch_image(void* batch, int32_t index, double* rgb, int64_t n_values);
int pppm_batch_scene_info(void* batch, int32_t index, int32_t* gather_mode, int32_t* photon_grid, int32_t* branch,
                          double* cost_ms);
/* The photon traversal's lanes in flight the tuner chose for a scene. */
int pppm_batch_scene_trace_keep(void* batch, int32_t index, int32_t* trace_keep);

/* ---- mesh files: replace Trimesh::load (sources/trimesh.cc:125-259) ---- */
typedef struct {
    int64_t n_vertices;
    float* vertices;  /* [n_vertices][3] */
    int64_t n_faces;
    int32_t* faces;   /* [n_faces][3] vertex indices, 0-based */
} pppm_mesh;
/* Reads a .ply (binary little / big endian or ASCII; x, y, z among any vertex
 * properties; face lists of any integer types, polygons fan-triangulated) or
 * .obj ("v" and "f" lines; 1-based, negative and slash-separated indices)
 * into host memory the library allocates; on failure err (may be NULL) gets
 * a message.  No device is touched. */
int pppm_mesh_load(const char* path, pppm_mesh* out, char* err, int32_t err_len);
void pppm_mesh_free(pppm_mesh* mesh);

/* ---- procedural scenes (no reference counterpart: the reference loads PLY
 * meshes, sources/trimesh.cc:125-217) ---- */
#define PPPM_SCENE_BOX 1        /* box + two icospheres */
#define PPPM_SCENE_BOX_CAUSTIC 2/* box + displaced icospheres + mirror + glass */
#define PPPM_SCENE_RANDOM 3     /* one scene of a batch: sized by param */
#define PPPM_SCENE_TWO_ROOMS 4  /* occluded light through a slit */
#define PPPM_SCENE_TWO_ROOMS_SSS 5

typedef struct {
    int64_t n_tris;
    float* tri_xyz;
    int32_t* bsdf_id;
    int32_t n_bsdf;
    pppm_bsdf* bsdfs;
    int32_t n_lights;
    pppm_light* lights;
    pppm_camera camera;
    double bbox_min[3];
    double bbox_max[3];
} pppm_scene_desc;

/* frequency: icosphere subdivision frequency (20 f^2 triangles per sphere);
 * 0 picks the documented default of the scene kind. */
int pppm_scenegen(int32_t kind, uint64_t seed, int32_t frequency, int32_t width, int32_t height,
                  pppm_scene_desc* out);
void pppm_scenegen_free(pppm_scene_desc* desc);

#ifdef __cplusplus
}
#endif
#endif
