// radix_sort.hip -- stable LSD radix sort, 8-bit digits, for gfx950.
//
// Per pass (reduce-then-scan):
//   1. k_sort_count:   per-tile digit histogram.  Each wavefront ranks its
//      64 keys of a round with an 8-ballot match (lanes holding the same
//      digit), so one LDS update per distinct digit instead of one atomic
//      per key: the nearly-sorted cell keys of a tile share few digits and
//      would otherwise serialise on a handful of LDS addresses.
//   2. k_sort_rowscan: one workgroup per digit scans that digit's row of
//      tile counts (exclusive) and records the digit total.
//   3. k_sort_scatter: re-ranks the tile (same match ranking, stable order
//      wave -> round -> lane), stages the tile in LDS in digit order and
//      writes each digit's run contiguously to its global offset.
#include "radix_sort.h"

namespace sph {

namespace {

__device__ __forceinline__ uint64_t lanes_below(int lane) {
    return lane == 0 ? 0ull : (~0ull >> (64 - lane));
}

// Lanes of the wavefront holding the same 8-bit digit as this lane,
// restricted to `active`.
__device__ __forceinline__ uint64_t match_digit(uint32_t d, uint64_t active) {
    uint64_t m = active;
#pragma unroll
    for (int b = 0; b < 8; ++b) {
        uint32_t bit = (d >> b) & 1u;
        uint64_t bal = __ballot(bit);
        m &= bit ? bal : ~bal;
    }
    return m;
}

// Exclusive scan of one value per thread across a 256-thread block.
__device__ __forceinline__ uint32_t block_scan_excl(uint32_t v, uint32_t* s_wave, uint32_t& total) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    uint32_t x = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        uint32_t y = __shfl_up(x, o, 64);
        if (lane >= o) x += y;
    }
    if (lane == 63) s_wave[wave] = x;
    __syncthreads();
    uint32_t off = 0;
    total = 0;
#pragma unroll
    for (int w = 0; w < kSortThreads / 64; ++w) {
        uint32_t t = s_wave[w];
        if (w < wave) off += t;
        total += t;
    }
    __syncthreads();
    return off + x - v;
}

// Ranks the keys of one tile.  Wave w owns the contiguous chunk
// [w*64*kSortItems, (w+1)*64*kSortItems) of the tile and walks it in
// kSortItems rounds of 64 consecutive keys.  On return s_cnt[w][d] holds
// the count of digit d in wave w's chunk and rank[r] the rank of key r among
// the equal digits of that chunk (stable).
__device__ __forceinline__ void rank_tile(const uint32_t (&k)[kSortItems], const bool (&valid)[kSortItems],
                                          int shift, uint32_t (*s_cnt)[kSortBins],
                                          uint32_t (&rank)[kSortItems]) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const uint64_t below = lanes_below(lane);
#pragma unroll
    for (int r = 0; r < kSortItems; ++r) {
        uint32_t d = (k[r] >> shift) & 0xFFu;
        uint64_t active = __ballot(valid[r]);
        uint64_t peers = match_digit(d, active);
        uint32_t base = 0;
        if (valid[r]) base = s_cnt[wave][d];
        int leader = __ffsll((unsigned long long)peers) - 1;
        if (valid[r] && lane == leader) s_cnt[wave][d] = base + (uint32_t)__popcll(peers);
        rank[r] = base + (uint32_t)__popcll(peers & below);
    }
}

__global__ __launch_bounds__(kSortThreads) void k_sort_count(const uint32_t* __restrict__ keys,
                                                             const uint32_t* __restrict__ n_dev, int shift,
                                                             uint32_t* __restrict__ hist, uint32_t max_tiles) {
    __shared__ uint32_t s_cnt[kSortThreads / 64][kSortBins];
    const uint32_t n = *n_dev;
    const uint32_t tile = blockIdx.x;
    const uint32_t t0 = tile * (uint32_t)kSortTile;
    if (t0 >= n) return;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (int i = threadIdx.x; i < (kSortThreads / 64) * kSortBins; i += kSortThreads)
        (&s_cnt[0][0])[i] = 0;
    __syncthreads();
    uint32_t k[kSortItems], rank[kSortItems];
    bool valid[kSortItems];
#pragma unroll
    for (int r = 0; r < kSortItems; ++r) {
        uint32_t idx = t0 + (uint32_t)(wave * 64 * kSortItems + r * 64 + lane);
        valid[r] = idx < n;
        k[r] = valid[r] ? keys[idx] : 0u;
    }
    rank_tile(k, valid, shift, s_cnt, rank);
    __syncthreads();
    const int d = threadIdx.x;
    uint32_t c = 0;
#pragma unroll
    for (int w = 0; w < kSortThreads / 64; ++w) c += s_cnt[w][d];
    hist[(size_t)d * max_tiles + tile] = c;
}

__global__ __launch_bounds__(kSortThreads) void k_sort_rowscan(uint32_t* __restrict__ hist,
                                                               uint32_t* __restrict__ totals,
                                                               const uint32_t* __restrict__ n_dev,
                                                               uint32_t max_tiles) {
    __shared__ uint32_t s_wave[kSortThreads / 64];
    const uint32_t n = *n_dev;
    const uint32_t tiles = (n + kSortTile - 1) / kSortTile;
    uint32_t* row = hist + (size_t)blockIdx.x * max_tiles;
    uint32_t carry = 0;
    for (uint32_t b = 0; b < tiles; b += kSortThreads) {
        uint32_t t = b + threadIdx.x;
        uint32_t v = t < tiles ? row[t] : 0u;
        uint32_t tot;
        uint32_t ex = block_scan_excl(v, s_wave, tot);
        if (t < tiles) row[t] = carry + ex;
        carry += tot;
    }
    if (threadIdx.x == 0) totals[blockIdx.x] = carry;
}

__global__ __launch_bounds__(kSortThreads) void k_sort_scatter(
    const uint32_t* __restrict__ keys_in, const uint32_t* __restrict__ vals_in, uint32_t* __restrict__ keys_out,
    uint32_t* __restrict__ vals_out, const uint32_t* __restrict__ n_dev, int shift,
    const uint32_t* __restrict__ hist, const uint32_t* __restrict__ totals, uint32_t max_tiles) {
    __shared__ uint32_t s_cnt[kSortThreads / 64][kSortBins];
    __shared__ uint32_t s_start[kSortBins];
    __shared__ uint32_t s_gbase[kSortBins];
    __shared__ uint32_t s_wave[kSortThreads / 64];
    __shared__ uint32_t s_key[kSortTile];
    __shared__ uint32_t s_val[kSortTile];
    const uint32_t n = *n_dev;
    const uint32_t tile = blockIdx.x;
    const uint32_t t0 = tile * (uint32_t)kSortTile;
    if (t0 >= n) return;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (int i = threadIdx.x; i < (kSortThreads / 64) * kSortBins; i += kSortThreads)
        (&s_cnt[0][0])[i] = 0;
    __syncthreads();
    uint32_t k[kSortItems], v[kSortItems], rank[kSortItems];
    bool valid[kSortItems];
#pragma unroll
    for (int r = 0; r < kSortItems; ++r) {
        uint32_t idx = t0 + (uint32_t)(wave * 64 * kSortItems + r * 64 + lane);
        valid[r] = idx < n;
        k[r] = valid[r] ? keys_in[idx] : 0u;
        v[r] = valid[r] ? (vals_in ? vals_in[idx] : idx) : 0u;
    }
    rank_tile(k, valid, shift, s_cnt, rank);
    __syncthreads();
    // Per digit: offsets of each wave's run inside the tile, tile total,
    // tile-local start of the digit, and the digit's global base.
    const int d = threadIdx.x;
    uint32_t tot = 0;
#pragma unroll
    for (int w = 0; w < kSortThreads / 64; ++w) {
        uint32_t c = s_cnt[w][d];
        s_cnt[w][d] = tot;
        tot += c;
    }
    uint32_t dummy;
    uint32_t start = block_scan_excl(tot, s_wave, dummy);
    s_start[d] = start;
    uint32_t gt = totals[d];
    uint32_t gbase = block_scan_excl(gt, s_wave, dummy);
    s_gbase[d] = gbase + hist[(size_t)d * max_tiles + tile];
    __syncthreads();
#pragma unroll
    for (int r = 0; r < kSortItems; ++r) {
        if (!valid[r]) continue;
        uint32_t dg = (k[r] >> shift) & 0xFFu;
        uint32_t lp = s_start[dg] + s_cnt[wave][dg] + rank[r];
        s_key[lp] = k[r];
        s_val[lp] = v[r];
    }
    __syncthreads();
    const uint32_t cnt = min((uint32_t)kSortTile, n - t0);
    for (uint32_t i = threadIdx.x; i < cnt; i += kSortThreads) {
        uint32_t key = s_key[i];
        uint32_t dg = (key >> shift) & 0xFFu;
        uint32_t g = s_gbase[dg] + (i - s_start[dg]);
        keys_out[g] = key;
        vals_out[g] = s_val[i];
    }
}

}  // namespace

hipError_t RadixSort::init(uint32_t cap) {
    release();
    capacity = cap < 1 ? 1 : cap;
    max_tiles = (capacity + kSortTile - 1) / kSortTile;
    hipError_t e;
    if ((e = hipMalloc(&key_alt, sizeof(uint32_t) * capacity)) != hipSuccess) return e;
    if ((e = hipMalloc(&val_alt, sizeof(uint32_t) * capacity)) != hipSuccess) return e;
    if ((e = hipMalloc(&hist, sizeof(uint32_t) * kSortBins * (size_t)max_tiles)) != hipSuccess) return e;
    if ((e = hipMalloc(&totals, sizeof(uint32_t) * kSortBins)) != hipSuccess) return e;
    return hipSuccess;
}

void RadixSort::release() {
    if (key_alt) (void)hipFree(key_alt);
    if (val_alt) (void)hipFree(val_alt);
    if (hist) (void)hipFree(hist);
    if (totals) (void)hipFree(totals);
    key_alt = val_alt = hist = totals = nullptr;
}

size_t RadixSort::bytes() const {
    return sizeof(uint32_t) * (2 * (size_t)capacity + kSortBins * (size_t)max_tiles + kSortBins);
}

void RadixSort::sort(hipStream_t st, const uint32_t* keys_in, const uint32_t* vals_in, uint32_t* keys_out,
                     uint32_t* vals_out, const uint32_t* n_dev, int bits) {
    int passes = (bits + 7) / 8;
    if (passes < 1) passes = 1;
    const uint32_t* kin = keys_in;
    const uint32_t* vin = vals_in;
    for (int p = 0; p < passes; ++p) {
        // The last pass lands in the output buffers.
        bool to_out = ((passes - 1 - p) % 2) == 0;
        uint32_t* kout = to_out ? keys_out : key_alt;
        uint32_t* vout = to_out ? vals_out : val_alt;
        int shift = 8 * p;
        hipLaunchKernelGGL(k_sort_count, dim3(max_tiles), dim3(kSortThreads), 0, st, kin, n_dev, shift, hist,
                           max_tiles);
        hipLaunchKernelGGL(k_sort_rowscan, dim3(kSortBins), dim3(kSortThreads), 0, st, hist, totals, n_dev,
                           max_tiles);
        hipLaunchKernelGGL(k_sort_scatter, dim3(max_tiles), dim3(kSortThreads), 0, st, kin, vin, kout, vout,
                           n_dev, shift, hist, totals, max_tiles);
        kin = kout;
        vin = vout;
    }
}

}  // namespace sph
