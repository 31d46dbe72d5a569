// cell_sort.hip -- counting sort by cell key (see cell_sort.h).
//
//   k_cs_count    one atomic per run of equal keys in a wavefront: the
//                 cells' counts and each particle's arrival rank in its cell;
//                 flags the 4096-cell blocks that hold a key
//   k_cs_reduce   totals of 4096-cell blocks (16 counts per thread, read as
//                 four 16-byte vectors); an unflagged block is empty and
//                 reads nothing (a fluid filling a tenth of the domain's
//                 cells touches about a tenth of the blocks)
//   k_cs_ranges   cell starts = offset of the block (sum of the totals of
//                 the blocks before it, read by the whole workgroup) + scan
//                 inside the block; writes range[k] of every occupied cell
//                 and clears its count and the block's flag
//   k_cs_scatter  slot = start of the cell + arrival rank
// (Block totals accumulated by atomics in k_cs_count were measured slower:
// neighbouring wavefronts add to the same total and serialise.)
// Two streaming reads of the count array (4 bytes per cell) replace the
// three count / scan / scatter passes of an LSD radix sort over the keys and
// the separate cell-range pass.  The arrival order inside a cell is not
// deterministic; the reorder pass ranks each cell by original id, so the
// sorted state is.
#include "cell_sort.h"

namespace sph {

namespace {

constexpr int kCsThreads = 256;
constexpr int kCsPerThread = 16;
constexpr int kCsBlockCells = kCsThreads * kCsPerThread;   // 4096

// The keys arrive nearly sorted (the storage order is the last step's cell
// order), so a wavefront holds a few runs of equal keys: the first lane of
// each run adds the run's length with one atomic and hands the old count to
// the run, whose lanes rank themselves by their offset in it.  Equal keys
// in separate runs take separate atomics, which is still exact.
__global__ __launch_bounds__(kCsThreads) void k_cs_count(const uint32_t* __restrict__ keys, uint32_t n_host,
                                                         const uint32_t* __restrict__ n_dev,
                                                         uint32_t* __restrict__ cnt, uint32_t* __restrict__ rank,
                                                         uint32_t* __restrict__ used,
                                                         const uint32_t* __restrict__ prev_skey, uint32_t prev_n,
                                                         int2* __restrict__ range) {
    const uint32_t n = n_dev ? *n_dev : n_host;
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    // invalidate the cells of the previous sort (its sorted keys, prev_n of
    // them) before k_cs_ranges writes this sort's cells
    if (prev_skey && i < prev_n) {
        const uint32_t pk = prev_skey[i];
        if (i == 0 || prev_skey[i - 1] != pk) range[pk] = make_int2(-1, -1);
    }
    if ((i & ~63u) >= n) return;   // the whole wavefront lies past the end
    const int lane = threadIdx.x & 63;
    const bool valid = i < n;
    const uint32_t k = valid ? keys[i] : 0xFFFFFFFFu;
    const uint32_t kp = __shfl_up(k, 1, 64);
    const bool head = valid && (lane == 0 || kp != k);
    const uint64_t heads = __ballot(head);
    const uint64_t vmask = __ballot(valid);   // a prefix of the lanes
    uint32_t base = 0;
    if (head) {
        const uint64_t after = lane == 63 ? 0ull : (heads >> (lane + 1));
        const int next = after ? lane + __ffsll((unsigned long long)after) : __popcll(vmask);
        base = atomicAdd(&cnt[k], (uint32_t)(next - lane));
        used[k / (uint32_t)kCsBlockCells] = 1u;
    }

    // the run's head: highest head lane at or below this lane
    const uint64_t upto = heads & (lane == 63 ? ~0ull : ((2ull << lane) - 1));
    const int hl = upto ? 63 - __clzll((long long)upto) : 0;
    const uint32_t b = __shfl(base, hl, 64);
    if (valid) rank[i] = b + (uint32_t)(lane - hl);
}

// The 16 counts of this thread (cells base .. base + 15, zero past the end).
__device__ __forceinline__ void load16(const uint32_t* __restrict__ cnt, uint32_t cells, uint32_t base,
                                       uint32_t (&c)[kCsPerThread]) {
    if (base + kCsPerThread <= cells) {
        const uint4* p = reinterpret_cast<const uint4*>(cnt + base);
#pragma unroll
        for (int v = 0; v < kCsPerThread / 4; ++v) {
            const uint4 q = p[v];
            c[4 * v] = q.x;
            c[4 * v + 1] = q.y;
            c[4 * v + 2] = q.z;
            c[4 * v + 3] = q.w;
        }
    } else {
#pragma unroll
        for (int j = 0; j < kCsPerThread; ++j) c[j] = base + j < cells ? cnt[base + j] : 0u;
    }
}

// Exclusive scan of one value per thread across the block; total in *total.
__device__ __forceinline__ uint32_t block_scan(uint32_t v, uint32_t* s_wave, uint32_t* total) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    uint32_t x = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t y = __shfl_up(x, o, 64);
        if (lane >= o) x += y;
    }
    if (lane == 63) s_wave[wave] = x;
    __syncthreads();
    uint32_t off = 0, tot = 0;
#pragma unroll
    for (int w = 0; w < kCsThreads / 64; ++w) {
        const uint32_t t = s_wave[w];
        if (w < wave) off += t;
        tot += t;
    }
    __syncthreads();
    *total = tot;
    return off + x - v;
}

__global__ __launch_bounds__(kCsThreads) void k_cs_reduce(const uint32_t* __restrict__ cnt, uint32_t cells,
                                                          const uint32_t* __restrict__ used,
                                                          uint32_t* __restrict__ bsum) {
    __shared__ uint32_t s_wave[kCsThreads / 64];
    if (!used[blockIdx.x]) {   // the same for the whole workgroup
        if (threadIdx.x == 0) bsum[blockIdx.x] = 0u;
        return;
    }
    const uint32_t base = blockIdx.x * (uint32_t)kCsBlockCells + threadIdx.x * (uint32_t)kCsPerThread;
    uint32_t c[kCsPerThread];
    load16(cnt, cells, base, c);
    uint32_t s = 0;
#pragma unroll
    for (int j = 0; j < kCsPerThread; ++j) s += c[j];
    uint32_t tot;
    (void)block_scan(s, s_wave, &tot);
    if (threadIdx.x == 0) bsum[blockIdx.x] = tot;
}

__global__ __launch_bounds__(kCsThreads) void k_cs_ranges(uint32_t* __restrict__ cnt, uint32_t cells,
                                                          const uint32_t* __restrict__ bsum,
                                                          uint32_t* __restrict__ used,
                                                          int2* __restrict__ range) {
    __shared__ uint32_t s_wave[kCsThreads / 64];
    // every thread reads the flag before the scans' barriers; thread 0
    // clears it after them
    if (!used[blockIdx.x]) return;   // no occupied cell in the block
    const uint32_t base = blockIdx.x * (uint32_t)kCsBlockCells + threadIdx.x * (uint32_t)kCsPerThread;
    uint32_t c[kCsPerThread];
    load16(cnt, cells, base, c);
    // offset of this block: the totals of the blocks before it
    uint32_t before = 0;
    for (uint32_t j = threadIdx.x; j < blockIdx.x; j += kCsThreads) before += bsum[j];
    uint32_t off;
    (void)block_scan(before, s_wave, &off);
    uint32_t s = 0;
#pragma unroll
    for (int j = 0; j < kCsPerThread; ++j) s += c[j];
    uint32_t tot;
    uint32_t start = off + block_scan(s, s_wave, &tot);
#pragma unroll
    for (int j = 0; j < kCsPerThread; ++j) {
        if (c[j]) {
            range[base + j] = make_int2((int)start, (int)(start + c[j]));
            cnt[base + j] = 0u;
        }
        start += c[j];
    }
    if (threadIdx.x == 0) used[blockIdx.x] = 0u;
}

__global__ __launch_bounds__(kCsThreads) void k_cs_scatter(const uint32_t* __restrict__ keys, uint32_t n_host,
                                                           const uint32_t* __restrict__ n_dev,
                                                           const uint32_t* __restrict__ rank,
                                                           const int2* __restrict__ range,
                                                           uint32_t* __restrict__ skey, uint32_t* __restrict__ perm) {
    const uint32_t n = n_dev ? *n_dev : n_host;
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t k = keys[i];
    const uint32_t d = (uint32_t)range[k].x + rank[i];
    skey[d] = k;
    perm[d] = i;
}

inline unsigned grid_for(uint32_t n) { return (n + kCsThreads - 1) / kCsThreads; }

}  // namespace

hipError_t CellSort::init(uint32_t ncells, uint32_t capacity) {
    release();
    cells = ncells;
    blocks = (cells + kCsBlockCells - 1) / kCsBlockCells;
    hipError_t e;
    // padded to whole blocks so that the vector loads of the last block stay
    // inside the allocation
    const size_t padded = (size_t)blocks * kCsBlockCells;
    if ((e = hipMalloc(&cnt, sizeof(uint32_t) * (padded ? padded : 1))) != hipSuccess) return e;
    if ((e = hipMalloc(&rank, sizeof(uint32_t) * (capacity ? capacity : 1))) != hipSuccess) return e;
    if ((e = hipMalloc(&bsum, sizeof(uint32_t) * (blocks ? blocks : 1))) != hipSuccess) return e;
    if ((e = hipMalloc(&used, sizeof(uint32_t) * (blocks ? blocks : 1))) != hipSuccess) return e;
    if ((e = hipMemset(cnt, 0, sizeof(uint32_t) * (padded ? padded : 1))) != hipSuccess) return e;
    if ((e = hipMemset(used, 0, sizeof(uint32_t) * (blocks ? blocks : 1))) != hipSuccess) return e;
    bytes = sizeof(uint32_t) * (padded + capacity + 2 * blocks);
    return hipDeviceSynchronize();
}

void CellSort::release() {
    if (cnt) (void)hipFree(cnt);
    if (rank) (void)hipFree(rank);
    if (bsum) (void)hipFree(bsum);
    if (used) (void)hipFree(used);
    cnt = rank = bsum = used = nullptr;
    bytes = 0;
}

void CellSort::sort(hipStream_t st, const uint32_t* keys, uint32_t n, uint32_t* skey, uint32_t* perm,
                    int2* range, bool clear_prev) {
    if (n == 0) return;
    launch(st, keys, n, nullptr, n, skey, perm, range, clear_prev ? skey : nullptr, n);
}

void CellSort::sort(hipStream_t st, const uint32_t* keys, const uint32_t* n_dev, uint32_t capacity, uint32_t* skey,
                    uint32_t* perm, int2* range) {
    if (capacity == 0) return;
    launch(st, keys, 0, n_dev, capacity, skey, perm, range, nullptr, 0);
}

void CellSort::launch(hipStream_t st, const uint32_t* keys, uint32_t n, const uint32_t* n_dev, uint32_t grid_n,
                      uint32_t* skey, uint32_t* perm, int2* range, const uint32_t* prev_skey, uint32_t prev_n) {
    hipLaunchKernelGGL(k_cs_count, dim3(grid_for(grid_n)), dim3(kCsThreads), 0, st, keys, n, n_dev, cnt, rank,
                       used, prev_skey, prev_n, range);
    hipLaunchKernelGGL(k_cs_reduce, dim3(blocks), dim3(kCsThreads), 0, st, cnt, cells, used, bsum);
    hipLaunchKernelGGL(k_cs_ranges, dim3(blocks), dim3(kCsThreads), 0, st, cnt, cells, bsum, used, range);
    hipLaunchKernelGGL(k_cs_scatter, dim3(grid_for(grid_n)), dim3(kCsThreads), 0, st, keys, n, n_dev, rank, range,
                       skey, perm);
}

}  // namespace sph
