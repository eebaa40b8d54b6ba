// internal.h -- host-side declarations shared by the translation units of
// libpppm_hip.so (context, device buffers, kernel launchers).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

#include "pppm_device.h"

// ---------------------------------------------------------------------------
// Stable LSD radix sort of (u32 key, u32 value) pairs, 8-bit digits.
// ---------------------------------------------------------------------------
#define RADIX_BLOCK 256
#define RADIX_ITEMS 8
#define RADIX_TILE (RADIX_BLOCK * RADIX_ITEMS)

inline int64_t radix_blocks(int64_t n) { return (n + RADIX_TILE - 1) / RADIX_TILE; }
// words of the histogram workspace for n keys (256 per tile + 256 totals)
inline int64_t radix_hist_words(int64_t n) { return 256 * (radix_blocks(n) + 1) + 256; }

// Sorts keys[0..n) by bits [0, bits), carrying vals.  Ping-pongs with the
// *_alt buffers; returns true if the result ended in the *_alt buffers.
bool radix_sort_pairs(hipStream_t s, uint32_t* keys, uint32_t* vals, uint32_t* keys_alt,
                      uint32_t* vals_alt, uint32_t* hist, int64_t n, int bits);

// ---------------------------------------------------------------------------
// Hash grid kernels (hash_grid.hip)
// ---------------------------------------------------------------------------
// Pass constants in device memory: a pass enqueued without host
// synchronisation (batched, graph-captured loops) reads its pass index and
// radius here, and k_pass_advance moves them to the next pass.
struct PassDev {
    double radius;   // r_i
    double scale;    // 1 / (2 r_i)
    uint32_t pass;   // i, 1-based
    uint32_t overflow;  // set when a pass stored more photons than the buffers hold
};

struct GridParams {
    double bmin[3];
    double scale;
    uint32_t size;
    // bits per axis of a cell-local Morton code below the bucket in the sort
    // key (0: records of a bucket in insertion order, the reference's layout)
    int32_t sub_bits = 0;
    const PassDev* pd = nullptr;  // when set, the scale is read from it
};
__device__ __forceinline__ double grid_scale(const GridParams& g) { return g.pd ? g.pd->scale : g.scale; }
// keys[i] = hash(cell(pos[perm ? perm[i] : i])), vals[i] = i; with a device
// count, slots i >= min(*count, n) get key `sentinel` (>= size)
void launch_point_keys(hipStream_t s, const float4* pos4, const float* pos3, const uint32_t* perm,
                       int64_t n, GridParams g, uint32_t* keys, uint32_t* vals,
                       const unsigned long long* count = nullptr, uint32_t sentinel = 0xFFFFFFFFu);
// Origin order without a sort (see k_origin_point_keys): start[i] = number of
// records of photons 0..i-1, from the per-photon counts the photon pass wrote
// (tile_sums: origin_tiles(photons) words of scratch) ...
int64_t origin_tiles(int64_t photons);
void launch_origin_starts(hipStream_t s, const uint32_t* count, int64_t photons, uint32_t* tile_sums,
                          uint32_t* start);
// ... then every record slot j goes to p = start[photon] + rank[j] with its
// grid key: keys[p], vals[p] = p, perm[p] = j (with a device count, unused
// slots and positions keep their place with key `sentinel`)
void launch_origin_point_keys(hipStream_t s, const float4* pos, const float4* flux, const uint8_t* rank,
                              const uint32_t* start, int64_t n, const unsigned long long* count, GridParams g,
                              uint32_t sentinel, uint32_t* keys, uint32_t* vals, uint32_t* perm, int stride = 1);
// off[h] = number of sorted keys whose bucket (key >> shift) is < h, for h = 0..size
// (scratch: at least (size + 1) / 4096 + 1 free words, or null)
void launch_bucket_offsets(hipStream_t s, const uint32_t* sorted_keys, int64_t n, uint32_t size,
                           uint32_t* off, int shift = 0, uint32_t* scratch = nullptr, int64_t scratch_words = 0);
// box insertion: per-item pair counts, then pairs in z, y, x order
void launch_box_counts(hipStream_t s, const double* bmin_items, const double* bmax_items,
                       int64_t n, GridParams g, uint64_t* counts);
void launch_box_pairs(hipStream_t s, const double* bmin_items, const double* bmax_items,
                      int64_t n, GridParams g, const uint64_t* offsets, uint32_t* keys,
                      uint32_t* vals);
void launch_exclusive_scan_u64(hipStream_t s, uint64_t* data, int64_t n, uint64_t* block_sums);

// ---------------------------------------------------------------------------
// BVH (bvh.hip)
// ---------------------------------------------------------------------------
struct BvhBuild {
    // outputs
    Node4* nodes;
    float4* tris4;
    int32_t root;
    int64_t n_nodes, n_leaves;
    int32_t depth;
    int32_t builder_used;  // BVH_BUILD_LBVH or BVH_BUILD_PLOC (PLOC falls back when its rounds run long)
};
// Binary trees the 4-wide BVH is collapsed from.
#define BVH_BUILD_LBVH 1  // radix tree over the Morton codes (Karras 2012)
#define BVH_BUILD_PLOC 2  // locally-ordered clustering on surface area (Meister and Bittner 2018)
// What PPPM_BVH_DEFAULT resolves to.
#ifndef BVH_BUILDER_DEFAULT
#define BVH_BUILDER_DEFAULT PPPM_BVH_PLOC
#endif
// Builds the 4-wide BVH of n triangles (tri9: [n][9] floats on the device)
// from the binary tree `builder` names.  Allocates nodes / tris4 with
// hipMalloc (caller frees).  Returns an error message, empty on success.
std::string bvh_build(hipStream_t s, const float* tri9, int64_t n, int builder, BvhBuild* out);
void launch_intersect(hipStream_t s, SceneDev sc, const pppm_ray* rays, int64_t n, pppm_hit* hits);

// ---------------------------------------------------------------------------
// Render kernels (render.hip)
// ---------------------------------------------------------------------------
struct Counters {
    unsigned long long photons_stored;
    unsigned long long photon_segments;
    unsigned long long eye_segments;
    unsigned long long hitpoints;
    unsigned long long gather_buckets;
    unsigned long long gather_candidates;
    unsigned long long gather_inradius;
    unsigned long long stack_overflow;
    unsigned long long proposal_records;
    unsigned long long irr_stored;   // photons stored on BSSRDF surfaces
    unsigned long long sss_hits;     // eye paths ending in the subsurface estimate
    unsigned long long photon_next;  // next photon index handed to a photon-pass lane
    unsigned long long gather_tested;  // photon records distance-tested by the gather
    unsigned long long gather_blocks;  // block boxes tested by the lanes gather
    // traversal work, counted by the counting kernels only
    unsigned long long photon_nodes, photon_tris, eye_nodes, eye_tris;
    unsigned long long gather_staged;  // records staged per wave by the blocks gather
};

// Persistent photon-pass blocks per CU: at least the blocks one CU holds at
// once (the kernel regenerates paths, so extra blocks find no work).
#define PHOTON_BLOCKS_PER_CU 8
// Eye-pass blocks per CU at most: one lane per pixel, so a 1024^2 image takes
// 4096 blocks, which the hardware hands out as earlier blocks finish.
#define EYE_BLOCKS_PER_CU 64

#ifndef PHOTON_TRACE_KEEP
#define PHOTON_TRACE_KEEP 16
#endif
struct PhotonPassArgs {
    SceneDev sc;
    int64_t n_photons;
    int64_t cap;
    uint64_t seed;
    uint32_t pass;
    uint32_t scene_id;
    int32_t max_bounces;
    float ray_eps;
    // record slot j at ph_pos[j * rec_stride], ph_flux[j * rec_stride]
    float4* ph_pos;   // xyz, oct normal bits
    float4* ph_flux;  // rgb, origin key bits (photon << 6 | bounce)
    int32_t rec_stride = 1;
    Counters* counters;
    const PassDev* pd = nullptr;  // when set, the pass index is read from it
    PropDev prop;     // emission proposal (side 0: plain cosine emission)
    // irradiance photons at BSSRDF surfaces (null: none stored)
    float4* irr_pos;  // xyz, oct normal bits
    float4* irr_flux; // rgb, origin key bits
    int64_t irr_cap;
    int32_t count_traversal = 0;  // 1: the kernel that also counts nodes and triangles
    // when set: rec_rank[slot] = the record's rank among its photon's stored
    // records, rec_count[photon index] = that photon's stored records
    uint8_t* rec_rank = nullptr;
    uint32_t* rec_count = nullptr;
    // the wave traces until at most this many of its lanes still have a ray
    // in flight, then regenerates paths (a launch-shape knob: the records do
    // not depend on it)
    int32_t trace_keep = PHOTON_TRACE_KEEP;
};
void launch_photon_pass(hipStream_t s, const PhotonPassArgs& a, int grid_limit);

// pixel id marking a pixel slot without a hit point
#define NO_HITPOINT 0xFFFFFFFFu

struct EyePassArgs {
    SceneDev sc;
    pppm_camera cam;
    float cam_basis[12];  // center, D, U * tanh * aspect, V * tanh
    uint64_t seed;
    uint32_t pass;
    uint32_t scene_id;
    int32_t max_bounces;
    float ray_eps;
    float4* hp_pos;  // xyz, oct normal bits
    float4* hp_w;    // weight rgb, pixel id bits (no hit point: the direct emission, id all ones)
    float* img_sum;   // direct emission, Kahan-added (null: left in the pixel slots)
    float* img_comp;
    Counters* counters;
    const PassDev* pd = nullptr;  // when set, the pass index is read from it
    // eye paths transmitted into a BSSRDF surface: (xyz, dipole index bits)
    // (weight rgb, pixel id bits), appended (null: none recorded)
    float4* sss_pos;
    float4* sss_w;
    int64_t sss_cap;
    int32_t count_traversal = 0;  // 1: the kernel that also counts nodes and triangles
    int32_t roulette = 1;         // 0: no eye Russian roulette (the Hachisuka method)
};
// grid_limit: at most this many blocks (a grid-stride loop covers the rest)
void launch_eye_pass(hipStream_t s, const EyePassArgs& a, int grid_limit);
// img += the direct emission stored in the pixel slots without hit point
// (an eye pass run with img_sum = null)
void launch_direct_add(hipStream_t s, const float4* hp_w, int64_t np, float* img_sum, float* img_comp);
// FP64 image average encoded per pixel (fmt 0: RGB8 after c^ginv, 1: RGBE, 2: FP32 RGB)
void launch_image_encode(hipStream_t s, const float* sum, const float* comp, const float* dsum, const float* dcomp,
                         int64_t npix, double inv, int fmt, double ginv, uint8_t* out8, float* outf);
// lane k < lanes (<= 64) Kahan-adds v[(i + k) % n], i = 0 .. n-1, into sum[k], comp[k]
void launch_kahan_probe(hipStream_t s, const float* v, int64_t n, int lanes, float* sum, float* comp);

#define GATHER_TILED PPPM_GATHER_TILED
#define GATHER_LANES PPPM_GATHER_LANES
#define GATHER_WAVE PPPM_GATHER_WAVE
#define GATHER_BLOCKS PPPM_GATHER_BLOCKS
// the lanes and blocks kernels read the compact two-word layout with block boxes
inline __host__ __device__ bool gather_compact(int32_t mode) { return mode == GATHER_LANES || mode == GATHER_BLOCKS; }
// PPPM_GATHER_AUTO resolves to the kernel measured fastest on MI355X
#define GATHER_AUTO_CHOICE GATHER_LANES

// photon records per block box of the lanes gather
#define GATHER_BLOCK 32

struct GatherArgs {
    int32_t mode;           // GATHER_LANES, GATHER_BLOCKS, GATHER_TILED or GATHER_WAVE
    // photons in bucket order.  Lanes: (x, y, z, oct normal) (Phi, key); the
    // others: (x, y, z, n_x) (n_y, n_z, Phi_r, Phi_g) (Phi_b, key, -, -)
    const float4* ph_a;
    const float4* ph_b;
    const float4* ph_c;
    const uint32_t* off;    // [S + 1]
    const float4* hp_pos;
    const float4* hp_w;
    const Counters* hp_count_src;  // device hit point count (counters->hitpoints) or null
    int64_t n_hit;                 // used when hp_count_src is null
    GridParams g;
    double radius;
    float* img_sum;   // Kahan-accumulated image (null: write out_rgb)
    float* img_comp;
    float* out_rgb;   // [pixel id][3] raw estimates (test hook)
    Counters* counters;
    uint32_t* used;   // lanes kernel: used[j] = 1 for accepted photon records j (may be null)
    // lanes kernel: per block of GATHER_BLOCK records the box (lo, hi) of their
    // positions; blocks whose box is farther than the radius are skipped
    const float4* boxes = nullptr;
};
__device__ __forceinline__ double gather_radius(const GatherArgs& a) { return a.g.pd ? a.g.pd->radius : a.radius; }
void launch_gather(hipStream_t s, const GatherArgs& a);
// used[sorted[j]] = 1 for each photon record j (bucket order) that the gather
// of a's hit points accepts (used is not cleared here)
void launch_mark_used(hipStream_t s, const GatherArgs& a, const uint32_t* sorted, uint32_t* used);
// used_order[sorted[j]] = used[j] for j < n
void launch_scatter_used(hipStream_t s, const uint32_t* used, const uint32_t* sorted, int64_t n, uint32_t* used_order);
// Morton keys of the half cells of the hit points; slots whose pixel id is
// NO_HITPOINT get key 0xFFFFFFFF and sort behind every hit point
// (radius from g.pd when set)
void launch_hit_keys(hipStream_t s, const float4* hp_pos, const float4* hp_w, int64_t cap, GridParams g,
                     double radius, uint32_t* keys, uint32_t* vals);
// gather layout of n photons: record k = photon perm[sorted[k]] (either may be
// null); C null selects the compact two-word layout of the lanes kernel, which
// then also writes the block boxes (as launch_gather_boxes) when boxes is set
// boxes[2 k], boxes[2 k + 1] = (lo, hi) of the positions of records
// [GATHER_BLOCK k, GATHER_BLOCK (k + 1)) of a (n records); lo.w holds the
// records' octahedral normal code when they all share it, else MIXED_NORMALS
// (0x8000 halves: a code oct_encode never makes)
void launch_gather_boxes(hipStream_t s, const float4* a, int64_t n, float4* boxes);
#define MIXED_NORMALS 0x80008000u
inline int64_t gather_box_count(int64_t n) { return (n + GATHER_BLOCK - 1) / GATHER_BLOCK; }
// Records the lanes kernel reads per wait, and the records a lanes-layout
// position array holds past its capacity (reads past a bucket are masked out).
#ifndef LANES_INFLIGHT
#define LANES_INFLIGHT 8
#endif
#define GATHER_PAD 8
void launch_gather_layout(hipStream_t s, const float4* pos, const float4* flux, const uint32_t* perm,
                          const uint32_t* sorted, int64_t n, float4* A, float4* B, float4* C,
                          float4* boxes = nullptr, int stride = 1);

// origin keys of the stored photons, and the permutation applied to the records;
// with a device count, slots i >= min(*count, n) get key `sentinel`
void launch_origin_keys(hipStream_t s, const float4* ph_flux, int64_t n, uint32_t* keys,
                        uint32_t* vals, const unsigned long long* count = nullptr,
                        uint32_t sentinel = 0xFFFFFFFFu);
// pd->pass += 1, radius = radii[pass - 1], scale = 1 / (2 radius); sets
// pd->overflow when counters->photons_stored exceeds cap (counters may be null)
void launch_pass_advance(hipStream_t s, PassDev* pd, const double* radii);
void launch_capacity_check(hipStream_t s, PassDev* pd, const Counters* counters, int64_t cap);
void launch_permute2(hipStream_t s, const float4* a_in, const float4* b_in, const uint32_t* perm,
                     const uint32_t* perm2, int64_t n, float4* a_out, float4* b_out, int stride = 1);
void launch_compose(hipStream_t s, const uint32_t* perm1, const uint32_t* perm2, int64_t n,
                    uint32_t* out);
void launch_zero_u32(hipStream_t s, void* p, int64_t words);
void launch_stream_copy(hipStream_t s, const float4* src, float4* dst, int64_t n);
void launch_pack_points(hipStream_t s, const float* pos3, const float* a3, const float* n3,
                        int64_t n, float4* pos4, float4* b4, int normals);

// ---------------------------------------------------------------------------
// Learned emission proposal (proposal.hip)
// ---------------------------------------------------------------------------
#define MLP_IN 32
#define MLP_HID 64
#define MLP_OUT 1024
#define MLP_SPLITS 8  // K chunks of the weight-gradient products (batch / 8: a multiple of 32)
// FP32 master parameters: W1 [64][32] b1, W2 [64][64] b2, W3 [64][64] b3, W4 [1024][64] b4
#define OW1 0
#define OB1 2048
#define OW2 2112
#define OB2 6208
#define OW3 6272
#define OB3 10368
#define OW4 10432
#define OB4 75968
#define MLP_NPARAM 76992
// BF16 mirror: W1, W2, W3, W4 and the transposes of W2, W3, W4
#define MW1 0
#define MW2 2048
#define MW3 6144
#define MW4 10240
#define MW2T 75776
#define MW3T 79872
#define MW4T 83968
#define MLP_NMIRROR 149504
// feature-major BF16 activations of a training batch, rows of [F][B]
#define ACT_X0 0
#define ACT_H1 32
#define ACT_H2 96
#define ACT_H3 160
#define ACT_D1 224
#define ACT_D2 288
#define ACT_D3 352
#define ACT_DZ 416
#define ACT_ROWS 1440

struct MlpArgs {
    const float* params;     // FP32 master (biases are read from here)
    const uint16_t* wb;      // BF16 mirror
    const float4* rec;       // training records (u, v, cell bits, weight)
    const uint32_t* idx;     // batch sample -> record (null: identity)
    const float* inv_wsum;   // 1 / sum of the batch weights
    const float2* uv;        // inference inputs
    int64_t n;               // samples (training: multiple of 256)
    float* prob;             // inference output [n][1024]
    uint16_t* act;           // training activations [ACT_ROWS][n]
    float* loss;             // weighted cross entropy (accumulated; may be null)
};
struct AdamArgs {
    float* params;
    float* m;
    float* v;
    const float* part;       // [MLP_SPLITS][part_stride] split-K gradient partials
    int64_t part_stride;
    uint16_t* wb;
    float* grad_out;         // reduced gradient (may be null)
    float lr, b1, b2, eps, c1, c2;  // c1 = 1 / (1 - b1^t), c2 = 1 / (1 - b2^t)
    int32_t apply;
};
struct ContribArgs {
    const float4* ph_pos;    // photon records as appended (slot j at [j * rec_stride])
    const float4* ph_flux;
    int32_t rec_stride;
    const uint32_t* perm;    // origin order
    int64_t n;
    const uint32_t* used;    // per record slot: a hit point's gather accepted it
    float* seg_c;            // [n]
    uint64_t* scan;          // [n]
    PropDev prop;
    uint64_t seed;
    uint32_t pass;
    uint32_t scene_id;
    float4* ring;
    int64_t ring_cap;
    int64_t head;
    Counters* counters;      // proposal_records = records of this pass
};
void launch_mlp(hipStream_t s, const MlpArgs& a, bool train);
void launch_mlp_wgrad(hipStream_t s, const uint16_t* act, int64_t B, float* part);
void launch_mlp_adam(hipStream_t s, const AdamArgs& a);
void launch_mlp_mirror(hipStream_t s, const float* params, uint16_t* wb);
void launch_mlp_prep(hipStream_t s, const float4* ring, int64_t n_valid, int64_t B, uint64_t seed, uint32_t step,
                     int identity, uint32_t* idx, float* inv_wsum);
void launch_prop_table(hipStream_t s, const float* prob, int32_t anchors, float* cdf);
void launch_contrib(hipStream_t s, const ContribArgs& a, uint64_t* block_sums);

// ---------------------------------------------------------------------------
// Subsurface scattering (subsurface.hip)
// ---------------------------------------------------------------------------
struct DipoleArgs {
    const float4* pts;        // surface points: xyz, -
    const float* E;           // irradiance [n_pts][3]
    int64_t n_pts;
    float area;               // surface area per point
    const DipoleDev* dip;     // per BSSRDF material
    const float4* hit_pos;    // xyz, dipole index bits
    const float4* hit_w;      // weight rgb, pixel id bits
    const Counters* count_src;  // sss_hits, or null: n_hit
    int64_t n_hit;
    float* img_sum;           // Kahan image (null: write out_rgb [hit][3])
    float* img_comp;
    float* out_rgb;
};
// L = W (1 - F_dr) / pi * area * sum_j R_d(|x - x_j|^2) E_j for every hit
void launch_dipole(hipStream_t s, const DipoleArgs& a);
// The octree cut of the same sum (dipole_cut.hip): nodes in preorder
struct CutDev {
    float4* a;              // position xyz (per pass), area
    float4* b;              // irradiance rgb (per pass), escape index | leaf bit 31
    const float4* lo;       // box min, first leaf point in perm (bits)
    const float4* hi;       // box max, leaf point count (bits; 0: interior node)
    const uint32_t* perm;   // points of the leaves
    int32_t n_nodes;
    float max_error;        // a node is taken whole when area / d^2 < max_error, x outside its box
};
// node values of `count` nodes of one level from E [point][3] and the deeper levels
void launch_cut_update(hipStream_t s, const CutDev& t, const int32_t* nodes, int64_t count, const float* E);
// launch_dipole's output with the sum walked over the cut
void launch_dipole_cut(hipStream_t s, const DipoleArgs& a, const CutDev& t);
// issue-rate probes (which 0: v_fma_f32 chains, 1: v_exp_f32 + v_fma_f32
// chains) over blocks x 256 threads writing out[blocks * 256]
void launch_issue_probe(hipStream_t s, float* out, int blocks, int which);
int64_t issue_probe_wave_instr(int which);  // wave instructions per wave
// stratified points on triangles: n points over the area CDF, uniform inside
// each triangle (host, deterministic)
