// noalloc.hip -- the hash grid build without allocation
// (pppm_hashgrid_build_noalloc).
//
// The same computation as pppm_hashgrid_build (stable radix sort of the
// points' Teschner buckets, then the bucket ranges), with the caller's
// device buffers in and out, and every intermediate -- sort keys and values,
// their ping-pong copies, the digit histograms, the bucket offsets -- in
// `scratch`, one fixed device array of this file.  No call on this path
// allocates device or host memory.
//
// Each GPU the library runs on holds its own copy of `scratch` until the
// process ends.  Its device address is looked up once per device at context
// creation (noalloc_setup); host code passes that address to the kernels,
// never the array's name.  Every call on a GPU shares that GPU's array:
// calls that overlap on one GPU race.
#include "ctx.h"

#include <mutex>

namespace {

constexpr int64_t kMaxPoints = PPPM_NOALLOC_MAX_POINTS;
constexpr int64_t kMaxBuckets = PPPM_NOALLOC_MAX_BUCKETS;
// radix_hist_words(kMaxPoints): 256 words per 2048-key tile + 256 totals + 256
constexpr int64_t kHistWords = 256 * ((kMaxPoints + RADIX_TILE - 1) / RADIX_TILE + 1) + 256;
// keys, values, keys_alt, values_alt | histograms | bucket offsets [S + 1]
constexpr int64_t kScratchWords = 4 * kMaxPoints + kHistWords + kMaxBuckets + 1;

}  // namespace

static __device__ __align__(256) uint32_t scratch[kScratchWords];

namespace {

std::mutex g_mu;
void* g_addr[64] = {};
bool g_done[64] = {};
hipError_t g_err[64] = {};

// sorted[j] = vals[j] for j < n; start[h] = off[h], end[h] = off[h + 1] for h < S
__global__ void k_noalloc_out(const uint32_t* __restrict__ vals, int64_t n, const uint32_t* __restrict__ off,
                              int64_t S, int32_t* __restrict__ sorted, int32_t* __restrict__ start,
                              int32_t* __restrict__ end) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) sorted[i] = (int32_t)vals[i];
    if (i < S) {
        start[i] = (int32_t)off[i];
        end[i] = (int32_t)off[i + 1];
    }
}

bool on_device(const void* p) {
    hipPointerAttribute_t a;
    if (hipPointerGetAttributes(&a, p) != hipSuccess) {
        (void)hipGetLastError();  // clear the sticky error of a plain host pointer
        return false;
    }
    return a.type == hipMemoryTypeDevice || a.type == hipMemoryTypeManaged;
}

}  // namespace

namespace pppmi {

// The device address of `scratch` on `device` (current), looked up once per
// device.
hipError_t noalloc_setup(int device, void** addr) {
    std::lock_guard<std::mutex> g(g_mu);
    if (device < 0 || device >= 64) return hipErrorInvalidDevice;
    if (!g_done[device]) {
        g_err[device] = hipGetSymbolAddress(&g_addr[device], HIP_SYMBOL(scratch));
        g_done[device] = true;
    }
    *addr = g_err[device] == hipSuccess ? g_addr[device] : nullptr;
    return g_err[device];
}

}  // namespace pppmi

using namespace pppmi;

extern "C" {

int pppm_hashgrid_build_noalloc(void* ctx_raw, const float* pos, int64_t n, int32_t hash_size, double hash_scale,
                                const double bbox_min[3], int32_t* sorted_idx, int32_t* cell_start,
                                int32_t* cell_end) {
    CHECK_CTX(ctx);
    if (!ctx->noalloc_scratch)
        return ctx->fail(PPPM_ERR_HIP, std::string("no scratch address on this device: ") +
                                           hipGetErrorString(ctx->noalloc_err));
    if (n < 0 || hash_size <= 0 || !bbox_min || !cell_start || !cell_end || (n > 0 && (!pos || !sorted_idx)))
        return ctx->fail(PPPM_ERR_INVALID, "bad hash grid arguments");
    if (n > kMaxPoints || hash_size > kMaxBuckets)
        return ctx->fail(PPPM_ERR_CAPACITY, "pppm_hashgrid_build_noalloc holds at most PPPM_NOALLOC_MAX_POINTS "
                                            "points and PPPM_NOALLOC_MAX_BUCKETS buckets");
    (void)hipSetDevice(ctx->device);
    if ((n > 0 && (!on_device(pos) || !on_device(sorted_idx))) || !on_device(cell_start) || !on_device(cell_end))
        return ctx->fail(PPPM_ERR_INVALID, "pppm_hashgrid_build_noalloc takes device buffers");
    uint32_t* w = static_cast<uint32_t*>(ctx->noalloc_scratch);
    uint32_t* keys = w;
    uint32_t* vals = w + kMaxPoints;
    uint32_t* keys_alt = w + 2 * kMaxPoints;
    uint32_t* vals_alt = w + 3 * kMaxPoints;
    uint32_t* hist = w + 4 * kMaxPoints;
    uint32_t* off = hist + kHistWords;
    GridParams g;
    for (int d = 0; d < 3; d++) g.bmin[d] = bbox_min[d];
    g.scale = hash_scale;
    g.size = (uint32_t)hash_size;
    hipStream_t s = ctx->stream;
    launch_point_keys(s, nullptr, pos, nullptr, n, g, keys, vals);
    const bool alt = radix_sort_pairs(s, keys, vals, keys_alt, vals_alt, hist, n, grid_bucket_bits(g.size, false));
    launch_bucket_offsets(s, alt ? keys_alt : keys, n, g.size, off, 0, hist, kHistWords);
    const int64_t m = std::max<int64_t>(n, hash_size);
    hipLaunchKernelGGL(k_noalloc_out, dim3((unsigned)((m + 255) / 256)), dim3(256), 0, s, alt ? vals_alt : vals, n,
                       off, (int64_t)hash_size, sorted_idx, cell_start, cell_end);
    TRY_HIP(ctx, hipGetLastError());
    TRY_HIP(ctx, hipStreamSynchronize(s));
    return PPPM_OK;
}

}  // extern "C"
