// hash_grid.hip -- GPU build of the reference HashGrid (sources/hash_grid_detail.h).
//
// Bucket of a point: h = ((ix*73856093) ^ (iy*19349663) ^ (iz*83492791)) mod S
// with i* = |trunc((x - bbox_min) * scale)| computed in FP64
// (hash_grid_detail.h:35-43, 60-62).  The int32 products of the reference
// wrap; uint32 products give the same bits.  A stable radix sort of
// (h, insertion index) then lays every bucket out in insertion order, which
// is exactly HashGrid<int>::_data[h] after the same add() calls.
#include "internal.h"

#pragma clang fp contract(off)

namespace {

__device__ inline int32_t cell_of(double x, double mn, double scale) {
    double b = (x - mn) * scale;
    int32_t c = (int32_t)b;  // truncation toward zero, as static_cast<int>
    return c < 0 ? -c : c;   // std::abs
}

__device__ inline uint32_t hash3(int32_t ix, int32_t iy, int32_t iz, uint32_t size) {
    uint32_t h = ((uint32_t)ix * 73856093u) ^ ((uint32_t)iy * 19349663u) ^ ((uint32_t)iz * 83492791u);
    return h % size;
}

// Sub-cell of x inside its cell along one axis, `bits` bits: the fraction of
// (x - min) * scale past its truncation.  Any value is valid (it orders records
// inside a bucket only); this one puts nearby records next to each other.
__device__ inline uint32_t sub_cell(double x, double mn, double scale, int bits) {
    const double b = (x - mn) * scale;
    const double f = fabs(b - trunc(b));
    const uint32_t k = (uint32_t)(f * (double)(1u << bits));
    return min(k, (1u << bits) - 1u);
}

__device__ inline uint32_t spread3(uint32_t v, int bits) {  // bit i -> bit 3 i
    uint32_t o = 0;
    for (int i = 0; i < bits; i++) o |= ((v >> i) & 1u) << (3 * i);
    return o;
}

// Grid sort key of a point: its bucket, then (sub_bits > 0) its cell-local
// Morton sub-cell below it.
__device__ inline uint32_t point_key(double x, double y, double z, const GridParams& g) {
    const double sc = grid_scale(g);
    uint32_t key = hash3(cell_of(x, g.bmin[0], sc), cell_of(y, g.bmin[1], sc), cell_of(z, g.bmin[2], sc), g.size);
    if (g.sub_bits > 0) {
        const int m = g.sub_bits;
        key = (key << (3 * m)) | spread3(sub_cell(x, g.bmin[0], sc, m), m) |
              (spread3(sub_cell(y, g.bmin[1], sc, m), m) << 1) | (spread3(sub_cell(z, g.bmin[2], sc, m), m) << 2);
    }
    return key;
}

__global__ void k_point_keys(const float4* __restrict__ pos4, const float* __restrict__ pos3,
                             const uint32_t* __restrict__ perm, int64_t n, GridParams g,
                             uint32_t* __restrict__ keys, uint32_t* __restrict__ vals,
                             const unsigned long long* __restrict__ count, uint32_t sentinel) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (count && i >= (int64_t)min((unsigned long long)n, *count)) {
        keys[i] = sentinel;
        vals[i] = (uint32_t)i;
        return;
    }
    int64_t j = perm ? (int64_t)perm[i] : i;
    double x, y, z;
    if (pos4) {
        float4 p = pos4[j];
        x = p.x; y = p.y; z = p.z;
    } else {
        x = pos3[3 * j]; y = pos3[3 * j + 1]; z = pos3[3 * j + 2];
    }
    keys[i] = point_key(x, y, z, g);
    vals[i] = (uint32_t)i;
}

// Exclusive scan of the per-photon record counts, in tiles of ORIGIN_TILE:
// the sum of every tile first, then each tile adds the sums of the tiles
// before it (a few thousand words, read from L2) and scans itself.
#define ORIGIN_TILE 4096
__global__ __launch_bounds__(256) void k_origin_tile_sums(const uint32_t* __restrict__ count, int64_t m,
                                                          uint32_t* __restrict__ tile_sums) {
    __shared__ uint32_t red[256];
    const int tid = threadIdx.x;
    const int64_t base = (int64_t)blockIdx.x * ORIGIN_TILE;
    uint32_t sum = 0;
    for (int r = 0; r < ORIGIN_TILE / 256; r++) {
        const int64_t i = base + (int64_t)r * 256 + tid;
        if (i < m) sum += count[i];
    }
    red[tid] = sum;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if (tid < off) red[tid] += red[tid + off];
        __syncthreads();
    }
    if (tid == 0) tile_sums[blockIdx.x] = red[0];
}

__global__ __launch_bounds__(256) void k_origin_starts(const uint32_t* __restrict__ count, int64_t m,
                                                       const uint32_t* __restrict__ tile_sums,
                                                       uint32_t* __restrict__ start) {
    __shared__ uint32_t red[256];
    __shared__ uint32_t s_scan[256];
    const int tid = threadIdx.x;
    constexpr int PER = ORIGIN_TILE / 256;  // consecutive counts per thread
    // the records of every earlier tile
    uint32_t before = 0;
    for (int64_t t = tid; t < (int64_t)blockIdx.x; t += 256) before += tile_sums[t];
    red[tid] = before;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if (tid < off) red[tid] += red[tid + off];
        __syncthreads();
    }
    const uint32_t tile_base = red[0];
    // this thread's run of PER counts, then a block scan of the run sums
    const int64_t base = (int64_t)blockIdx.x * ORIGIN_TILE + (int64_t)tid * PER;
    uint32_t c[PER];
    uint32_t run = 0;
#pragma unroll
    for (int k = 0; k < PER; k++) {
        c[k] = base + k < m ? count[base + k] : 0u;
        run += c[k];
    }
    s_scan[tid] = run;
    __syncthreads();
    for (int off = 1; off < 256; off <<= 1) {
        const uint32_t t = tid >= off ? s_scan[tid - off] : 0u;
        __syncthreads();
        s_scan[tid] += t;
        __syncthreads();
    }
    uint32_t acc = tile_base + s_scan[tid] - run;
#pragma unroll
    for (int k = 0; k < PER; k++) {
        if (base + k < m) start[base + k] = acc;
        acc += c[k];
    }
}

// Every position a sentinel at its own place: with a device count, a pass
// whose records overflowed the capacity leaves positions that no record
// reaches (the pass is reported as failed, but what follows must stay in
// bounds).
__global__ void k_origin_fill(int64_t n, uint32_t sentinel, uint32_t* __restrict__ keys,
                              uint32_t* __restrict__ vals, uint32_t* __restrict__ perm) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    keys[p] = sentinel;
    vals[p] = (uint32_t)p;
    perm[p] = (uint32_t)p;
}

// Photon records in origin order without a sort, and their grid keys: record
// slot j of photon i (key word i << 6 | bounce) holding rank r among that
// photon's stored records goes to position p = start[i] + r, which is its
// place in the stable sort by photon index, since a photon's records are
// appended in bounce order.  keys[p] = grid key, vals[p] = perm[p] = j: the
// grid sort then carries record slots, and the gather layout reads each
// record through one index instead of two.
// With a device count, slots j >= min(*count, n) keep their place with key
// `sentinel`; a position past n (records the capacity dropped: the pass is
// reported as failed) is not written, and the positions those records would
// have left are sentinels from k_origin_fill.
__global__ __launch_bounds__(256) void k_origin_point_keys(
    const float4* __restrict__ pos, const float4* __restrict__ flux, const uint8_t* __restrict__ rank,
    const uint32_t* __restrict__ start, int64_t n, const unsigned long long* __restrict__ count, GridParams g,
    uint32_t sentinel, uint32_t* __restrict__ keys, uint32_t* __restrict__ vals, uint32_t* __restrict__ perm,
    int stride) {
    const int64_t j = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (j >= n) return;
    if (count && j >= (int64_t)min((unsigned long long)n, *count)) {
        keys[j] = sentinel;
        vals[j] = (uint32_t)j;
        perm[j] = (uint32_t)j;
        return;
    }
    const uint32_t i = __float_as_uint(flux[j * stride].w) >> 6;
    const int64_t p = (int64_t)start[i] + rank[j];
    if (p >= n) return;
    const float4 q = pos[j * stride];
    keys[p] = point_key(q.x, q.y, q.z, g);
    vals[p] = (uint32_t)j;
    perm[p] = (uint32_t)j;
}

// off[h] = lower_bound(sorted buckets, h) for h in [0, size], the bucket of a
// key being key >> shift; each bucket boundary fills the empty buckets up to
// it.  Buckets >= size (sentinels of unused slots) may follow the real ones.
__global__ void k_bucket_offsets(const uint32_t* __restrict__ keys, int64_t n, uint32_t size, int shift,
                                 uint32_t* __restrict__ off) {
    int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j > n) return;
    const int64_t prev = (j == 0) ? -1 : (int64_t)(keys[j - 1] >> shift);
    const int64_t cur = (j == n) ? (int64_t)size : (int64_t)(keys[j] >> shift);
    if (j > 0 && j < n && cur == prev) return;  // inside a bucket
    int64_t lo = prev + 1;
    int64_t hi = cur;
    if (hi > (int64_t)size) hi = size;  // sentinel keys (>= size) close the last bucket
    for (int64_t h = lo; h <= hi; h++) off[h] = (uint32_t)j;
}

// Faster form for sparse grids: off = n everywhere, then off[b] = first record
// of each non-empty bucket b (and off[size] = the first sentinel slot), then a
// suffix minimum over off[0..size] fills every empty bucket with the start of
// the next non-empty one.
__global__ void k_fill_u32(uint32_t* __restrict__ p, int64_t m, uint32_t v) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < m) p[i] = v;
}

__global__ void k_mark_bucket_starts(const uint32_t* __restrict__ keys, int64_t n, uint32_t size, int shift,
                                     uint32_t* __restrict__ off) {
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    const uint32_t b = keys[j] >> shift;
    if (b >= size) {
        // sentinel slots (unused records) follow the real ones: the first
        // closes the grid
        if (j == 0 || (keys[j - 1] >> shift) < size) off[size] = (uint32_t)j;
        return;
    }
    if (j == 0 || (keys[j - 1] >> shift) != b) off[b] = (uint32_t)j;
}

#define SMIN_ITEMS 16
#define SMIN_TILE (256 * SMIN_ITEMS)

// Suffix minimum inside tiles of SMIN_TILE entries; tile minima to tmin.
__global__ __launch_bounds__(256) void k_suffix_min_tiles(uint32_t* __restrict__ a, int64_t m,
                                                          uint32_t* __restrict__ tmin) {
    __shared__ uint32_t s[256];
    const int tid = threadIdx.x;
    const int64_t base = (int64_t)blockIdx.x * SMIN_TILE + (int64_t)tid * SMIN_ITEMS;
    uint32_t v[SMIN_ITEMS];
    uint32_t run = 0xFFFFFFFFu;
#pragma unroll
    for (int k = SMIN_ITEMS - 1; k >= 0; k--) {
        const uint32_t x = base + k < m ? a[base + k] : 0xFFFFFFFFu;
        run = min(run, x);
        v[k] = run;
    }
    // suffix minimum of the thread minima, right to left
    s[tid] = run;
    __syncthreads();
    for (int off = 1; off < 256; off <<= 1) {
        const uint32_t t = tid + off < 256 ? s[tid + off] : 0xFFFFFFFFu;
        __syncthreads();
        s[tid] = min(s[tid], t);
        __syncthreads();
    }
    const uint32_t right = tid + 1 < 256 ? s[tid + 1] : 0xFFFFFFFFu;
#pragma unroll
    for (int k = 0; k < SMIN_ITEMS; k++)
        if (base + k < m) a[base + k] = min(v[k], right);
    if (tid == 0) tmin[blockIdx.x] = s[0];
}

// One workgroup: suffix minimum of the tile minima, in place.
__global__ __launch_bounds__(256) void k_suffix_min_top(uint32_t* __restrict__ tmin, int64_t nt) {
    __shared__ uint32_t s[256];
    __shared__ uint32_t carry;
    const int tid = threadIdx.x;
    if (tid == 0) carry = 0xFFFFFFFFu;
    __syncthreads();
    for (int64_t c = ((nt + 255) / 256 - 1) * 256; c >= 0; c -= 256) {
        const int64_t i = c + tid;
        s[tid] = i < nt ? tmin[i] : 0xFFFFFFFFu;
        __syncthreads();
        for (int off = 1; off < 256; off <<= 1) {
            const uint32_t t = tid + off < 256 ? s[tid + off] : 0xFFFFFFFFu;
            __syncthreads();
            s[tid] = min(s[tid], t);
            __syncthreads();
        }
        const uint32_t cv = carry;
        if (i < nt) tmin[i] = min(s[tid], cv);
        __syncthreads();
        if (tid == 0) carry = min(s[0], cv);
        __syncthreads();
    }
}

// a[i] = min(a[i], suffix minimum of the tiles to the right of i's tile)
__global__ void k_suffix_min_apply(uint32_t* __restrict__ a, int64_t m, const uint32_t* __restrict__ tmin,
                                   int64_t nt) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    const int64_t t = i / SMIN_TILE;
    if (t + 1 < nt) a[i] = min(a[i], tmin[t + 1]);
}

__device__ inline void box_cells(const double* bmin_items, const double* bmax_items, int64_t i,
                                 const GridParams& g, int32_t lo[3], int32_t hi[3]) {
    for (int d = 0; d < 3; d++) {
        lo[d] = cell_of(bmin_items[3 * i + d], g.bmin[d], g.scale);
        hi[d] = cell_of(bmax_items[3 * i + d], g.bmin[d], g.scale);
    }
}

__global__ void k_box_counts(const double* __restrict__ bmin_items,
                             const double* __restrict__ bmax_items, int64_t n, GridParams g,
                             uint64_t* __restrict__ counts) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int32_t lo[3], hi[3];
    box_cells(bmin_items, bmax_items, i, g, lo, hi);
    uint64_t c = 0;
    if (hi[0] >= lo[0] && hi[1] >= lo[1] && hi[2] >= lo[2])
        c = (uint64_t)(hi[0] - lo[0] + 1) * (uint64_t)(hi[1] - lo[1] + 1) * (uint64_t)(hi[2] - lo[2] + 1);
    counts[i] = c;
}

// hash_grid_detail.h:44-50: z outer, then y, then x, bounds inclusive
__global__ void k_box_pairs(const double* __restrict__ bmin_items,
                            const double* __restrict__ bmax_items, int64_t n, GridParams g,
                            const uint64_t* __restrict__ offsets, uint32_t* __restrict__ keys,
                            uint32_t* __restrict__ vals) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int32_t lo[3], hi[3];
    box_cells(bmin_items, bmax_items, i, g, lo, hi);
    uint64_t k = offsets[i];
    for (int32_t iz = lo[2]; iz <= hi[2]; iz++)
        for (int32_t iy = lo[1]; iy <= hi[1]; iy++)
            for (int32_t ix = lo[0]; ix <= hi[0]; ix++) {
                keys[k] = hash3(ix, iy, iz, g.size);
                vals[k] = (uint32_t)i;
                k++;
            }
}

// exclusive scan of u64, 1024 items per block, two levels
__global__ __launch_bounds__(256) void k_scan_u64_block(uint64_t* data, int64_t n,
                                                        uint64_t* block_sums) {
    __shared__ uint64_t s[256];
    const int tid = threadIdx.x;
    const int64_t base = (int64_t)blockIdx.x * 1024 + tid * 4;
    uint64_t v[4], sum = 0;
    for (int k = 0; k < 4; k++) {
        v[k] = base + k < n ? data[base + k] : 0;
        sum += v[k];
    }
    s[tid] = sum;
    __syncthreads();
    for (int off = 1; off < 256; off <<= 1) {
        uint64_t t = tid >= off ? s[tid - off] : 0;
        __syncthreads();
        s[tid] += t;
        __syncthreads();
    }
    uint64_t run = s[tid] - sum;
    for (int k = 0; k < 4; k++) {
        if (base + k < n) data[base + k] = run;
        run += v[k];
    }
    if (tid == 255) block_sums[blockIdx.x] = s[255];
}

__global__ void k_scan_u64_add(uint64_t* data, int64_t n, const uint64_t* block_prefix) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) data[i] += block_prefix[i / 1024];
}

}  // namespace

static inline unsigned grid1(int64_t n, int b) { return (unsigned)((n + b - 1) / b); }

void launch_point_keys(hipStream_t s, const float4* pos4, const float* pos3, const uint32_t* perm,
                       int64_t n, GridParams g, uint32_t* keys, uint32_t* vals, const unsigned long long* count,
                       uint32_t sentinel) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_point_keys, dim3(grid1(n, 256)), dim3(256), 0, s, pos4, pos3, perm, n, g,
                       keys, vals, count, sentinel);
}

int64_t origin_tiles(int64_t photons) { return (photons + ORIGIN_TILE - 1) / ORIGIN_TILE; }

void launch_origin_starts(hipStream_t s, const uint32_t* count, int64_t photons, uint32_t* tile_sums,
                          uint32_t* start) {
    if (photons <= 0) return;
    const int64_t nt = origin_tiles(photons);
    hipLaunchKernelGGL(k_origin_tile_sums, dim3((unsigned)nt), dim3(256), 0, s, count, photons, tile_sums);
    hipLaunchKernelGGL(k_origin_starts, dim3((unsigned)nt), dim3(256), 0, s, count, photons, tile_sums, start);
}

void launch_origin_point_keys(hipStream_t s, const float4* pos, const float4* flux, const uint8_t* rank,
                              const uint32_t* start, int64_t n, const unsigned long long* count, GridParams g,
                              uint32_t sentinel, uint32_t* keys, uint32_t* vals, uint32_t* perm, int stride) {
    if (n <= 0) return;
    if (count)
        hipLaunchKernelGGL(k_origin_fill, dim3(grid1(n, 256)), dim3(256), 0, s, n, sentinel, keys, vals, perm);
    hipLaunchKernelGGL(k_origin_point_keys, dim3(grid1(n, 256)), dim3(256), 0, s, pos, flux, rank, start, n, count,
                       g, sentinel, keys, vals, perm, stride);
}

void launch_bucket_offsets(hipStream_t s, const uint32_t* sorted_keys, int64_t n, uint32_t size,
                           uint32_t* off, int shift, uint32_t* scratch, int64_t scratch_words) {
    const int64_t m = (int64_t)size + 1;
    const int64_t nt = (m + SMIN_TILE - 1) / SMIN_TILE;
    if (!scratch || scratch_words < nt) {
        // one thread per bucket boundary fills the empty buckets before it
        hipLaunchKernelGGL(k_bucket_offsets, dim3(grid1(n + 1, 256)), dim3(256), 0, s, sorted_keys, n,
                           size, shift, off);
        return;
    }
    hipLaunchKernelGGL(k_fill_u32, dim3(grid1(m, 256)), dim3(256), 0, s, off, m, (uint32_t)n);
    if (n > 0)
        hipLaunchKernelGGL(k_mark_bucket_starts, dim3(grid1(n, 256)), dim3(256), 0, s, sorted_keys, n, size, shift,
                           off);
    hipLaunchKernelGGL(k_suffix_min_tiles, dim3((unsigned)nt), dim3(256), 0, s, off, m, scratch);
    hipLaunchKernelGGL(k_suffix_min_top, dim3(1), dim3(256), 0, s, scratch, nt);
    hipLaunchKernelGGL(k_suffix_min_apply, dim3(grid1(m, 256)), dim3(256), 0, s, off, m, scratch, nt);
}

void launch_box_counts(hipStream_t s, const double* bmin_items, const double* bmax_items,
                       int64_t n, GridParams g, uint64_t* counts) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_box_counts, dim3(grid1(n, 256)), dim3(256), 0, s, bmin_items, bmax_items,
                       n, g, counts);
}

void launch_box_pairs(hipStream_t s, const double* bmin_items, const double* bmax_items,
                      int64_t n, GridParams g, const uint64_t* offsets, uint32_t* keys,
                      uint32_t* vals) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_box_pairs, dim3(grid1(n, 256)), dim3(256), 0, s, bmin_items, bmax_items,
                       n, g, offsets, keys, vals);
}

// block_sums needs ceil(n/1024) + ceil(ceil(n/1024)/1024) + ... words
void launch_exclusive_scan_u64(hipStream_t s, uint64_t* data, int64_t n, uint64_t* block_sums) {
    if (n <= 0) return;
    int64_t nb = (n + 1023) / 1024;
    hipLaunchKernelGGL(k_scan_u64_block, dim3((unsigned)nb), dim3(256), 0, s, data, n, block_sums);
    if (nb > 1) {
        launch_exclusive_scan_u64(s, block_sums, nb, block_sums + nb);
        hipLaunchKernelGGL(k_scan_u64_add, dim3(grid1(n, 256)), dim3(256), 0, s, data, n,
                           block_sums);
    }
}
