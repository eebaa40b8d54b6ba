// radix.hip -- stable LSD radix sort of (u32 key, u32 value) pairs for gfx950.
//
// One pass per 8-bit digit, three kernels per pass:
//   1. digit histogram of each 2048-key tile (LDS atomics),
//   2. per-digit exclusive scan over tiles (one workgroup per digit),
//   3. stable scatter: each tile is ranked in 8 rounds of 256 consecutive keys;
//      inside a wave, keys with equal digits are matched with 8 ballots, waves
//      are ordered through a per-round LDS table, so the output order of equal
//      digits is the input order (the property the hash grid needs: a bucket
//      lists its items in insertion order).  The ranked tile is laid out in
//      LDS in digit order and written from there, so the stores of a wave fall
//      in runs of consecutive positions.
#include "internal.h"

namespace {

__global__ __launch_bounds__(RADIX_BLOCK) void k_digit_hist(const uint32_t* __restrict__ keys,
                                                            int64_t n, int shift,
                                                            uint32_t* __restrict__ hist,
                                                            int64_t nblocks) {
    __shared__ uint32_t h[256];
    const int tid = threadIdx.x;
    h[tid] = 0;
    __syncthreads();
    const int64_t base = (int64_t)blockIdx.x * RADIX_TILE;
#pragma unroll
    for (int r = 0; r < RADIX_ITEMS; r++) {
        int64_t i = base + (int64_t)r * RADIX_BLOCK + tid;
        if (i < n) atomicAdd(&h[(keys[i] >> shift) & 0xFFu], 1u);
    }
    __syncthreads();
    hist[(int64_t)tid * nblocks + blockIdx.x] = h[tid];
}

// One workgroup per digit: exclusive scan of hist[d][0..nblocks) in place;
// the digit total goes to totals[d].
__global__ __launch_bounds__(256) void k_digit_scan(uint32_t* __restrict__ hist, int64_t nblocks,
                                                    uint32_t* __restrict__ totals) {
    __shared__ uint32_t s[256];
    __shared__ uint32_t carry;
    const int tid = threadIdx.x;
    uint32_t* row = hist + (int64_t)blockIdx.x * nblocks;
    if (tid == 0) carry = 0;
    __syncthreads();
    for (int64_t c = 0; c < nblocks; c += 256) {
        int64_t i = c + tid;
        uint32_t v = i < nblocks ? row[i] : 0u;
        s[tid] = v;
        __syncthreads();
        for (int off = 1; off < 256; off <<= 1) {
            uint32_t t = tid >= off ? s[tid - off] : 0u;
            __syncthreads();
            s[tid] += t;
            __syncthreads();
        }
        uint32_t incl = s[tid];
        uint32_t c0 = carry;
        if (i < nblocks) row[i] = c0 + incl - v;
        __syncthreads();
        if (tid == 255) carry = c0 + incl;
        __syncthreads();
    }
    if (tid == 0) totals[blockIdx.x] = carry;
}

__global__ __launch_bounds__(RADIX_BLOCK) void k_scatter(
    const uint32_t* __restrict__ keys_in, const uint32_t* __restrict__ vals_in,
    uint32_t* __restrict__ keys_out, uint32_t* __restrict__ vals_out,
    const uint32_t* __restrict__ hist, const uint32_t* __restrict__ totals, int64_t n, int shift,
    int64_t nblocks) {
    constexpr int WAVES = RADIX_BLOCK / 64;
    __shared__ uint32_t s_base[256];    // global position of this tile's first key of each digit
    __shared__ uint32_t s_start[256];   // position in the tile of its first key of each digit
    __shared__ uint32_t s_run[256];
    __shared__ uint32_t s_cnt[WAVES][256];
    __shared__ uint32_t s_scan[256];
    __shared__ uint32_t s_key[RADIX_TILE];
    __shared__ uint32_t s_val[RADIX_TILE];
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;

    // digit bases: exclusive scan of the digit totals
    uint32_t tot = totals[tid];
    s_scan[tid] = tot;
    __syncthreads();
    for (int off = 1; off < 256; off <<= 1) {
        uint32_t t = tid >= off ? s_scan[tid - off] : 0u;
        __syncthreads();
        s_scan[tid] += t;
        __syncthreads();
    }
    s_base[tid] = (s_scan[tid] - tot) + hist[(int64_t)tid * nblocks + blockIdx.x];
    s_run[tid] = 0;

    // 1. stable rank of every key among the tile's keys of its digit: rounds of
    //    256 consecutive keys; inside a wave equal digits are matched with 8
    //    ballots, waves are ordered through a per-round LDS table
    const uint64_t lt_mask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    const int64_t base = (int64_t)blockIdx.x * RADIX_TILE;
    uint32_t key[RADIX_ITEMS], val[RADIX_ITEMS], dig[RADIX_ITEMS], idx[RADIX_ITEMS];
#pragma unroll
    for (int r = 0; r < RADIX_ITEMS; r++) {
        for (int w = 0; w < WAVES; w++) s_cnt[w][tid] = 0;
        __syncthreads();
        const int64_t i = base + (int64_t)r * RADIX_BLOCK + tid;
        const bool valid = i < n;
        key[r] = valid ? keys_in[i] : 0u;
        val[r] = valid ? vals_in[i] : 0u;
        const uint32_t d = (key[r] >> shift) & 0xFFu;
        dig[r] = valid ? d : 0xFFFFFFFFu;
        uint64_t peers = __ballot(valid);
#pragma unroll
        for (int b = 0; b < 8; b++) {
            bool bit = (d >> b) & 1u;
            uint64_t m = __ballot(bit);
            peers &= bit ? m : ~m;
        }
        const uint32_t rank = __popcll(peers & lt_mask);
        if (valid && (63 - __clzll(peers)) == lane) s_cnt[wave][d] = __popcll(peers);
        __syncthreads();
        {
            uint32_t run = s_run[tid];
#pragma unroll
            for (int w = 0; w < WAVES; w++) {
                uint32_t c = s_cnt[w][tid];
                s_cnt[w][tid] = run;
                run += c;
            }
            s_run[tid] = run;
        }
        __syncthreads();
        idx[r] = valid ? s_cnt[wave][d] + rank : 0u;
        __syncthreads();  // s_cnt is cleared by the next round
    }
    // 2. the tile in digit order in LDS: s_start = exclusive scan of the
    //    tile's digit counts (s_run)
    const uint32_t cnt = s_run[tid];
    s_scan[tid] = cnt;
    __syncthreads();
    for (int off = 1; off < 256; off <<= 1) {
        uint32_t t = tid >= off ? s_scan[tid - off] : 0u;
        __syncthreads();
        s_scan[tid] += t;
        __syncthreads();
    }
    s_start[tid] = s_scan[tid] - cnt;
    __syncthreads();
#pragma unroll
    for (int r = 0; r < RADIX_ITEMS; r++) {
        if (dig[r] != 0xFFFFFFFFu) {
            const uint32_t t = s_start[dig[r]] + idx[r];
            s_key[t] = key[r];
            s_val[t] = val[r];
        }
    }
    __syncthreads();
    // 3. write out in tile order: keys of one digit go to consecutive
    //    positions, so a wave's stores fall in runs
    const int64_t m = min((int64_t)RADIX_TILE, n - base);
#pragma unroll
    for (int r = 0; r < RADIX_ITEMS; r++) {
        const int t = r * RADIX_BLOCK + tid;
        if (t < m) {
            const uint32_t k = s_key[t];
            const uint32_t d = (k >> shift) & 0xFFu;
            const uint32_t pos = s_base[d] + (uint32_t)t - s_start[d];
            keys_out[pos] = k;
            vals_out[pos] = s_val[t];
        }
    }
}

}  // namespace

bool radix_sort_pairs(hipStream_t s, uint32_t* keys, uint32_t* vals, uint32_t* keys_alt,
                      uint32_t* vals_alt, uint32_t* hist, int64_t n, int bits) {
    if (n <= 1 || bits <= 0) return false;
    const int64_t nb = radix_blocks(n);
    uint32_t* totals = hist + 256 * (nb + 1);
    uint32_t *ki = keys, *vi = vals, *ko = keys_alt, *vo = vals_alt;
    bool alt = false;
    for (int shift = 0; shift < bits; shift += 8) {
        hipLaunchKernelGGL(k_digit_hist, dim3((unsigned)nb), dim3(RADIX_BLOCK), 0, s, ki, n, shift,
                           hist, nb);
        hipLaunchKernelGGL(k_digit_scan, dim3(256), dim3(256), 0, s, hist, nb, totals);
        hipLaunchKernelGGL(k_scatter, dim3((unsigned)nb), dim3(RADIX_BLOCK), 0, s, ki, vi, ko, vo,
                           hist, totals, n, shift, nb);
        uint32_t* t = ki; ki = ko; ko = t;
        t = vi; vi = vo; vo = t;
        alt = !alt;
    }
    return alt;
}
