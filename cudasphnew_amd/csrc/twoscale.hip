// twoscale.hip -- the two-scale sub-particle pass of AdvanceTwoScale
// (particle_simulation.cu:2145-2182), device-resident.
//
// Runs between the coarse density pass and the coarse force pass (the sub
// pass reads coarse positions, velocities and densities before integration,
// as the reference's order does; nothing feeds back into the coarse step):
//   state    shift_state + find_split_particles + find_boundary_particles
//            (cu:730-740, 932-1017); the boundary bit is pushed from the
//            split particles as in the reference, but into a separate final
//            state array: every writer of a byte stores the same value
//            (state | 2) instead of the reference's racy byte OR.  With the
//            hit-mask pair the push follows the coarse density pass's row
//            words (the split particle's neighbours within h) instead of
//            searching the cell table.
//   collect  check_split_boundary_default + 3 exclusive scans + collect_ids
//            (cu:1092-1162, 2509-2567) as one 3-way compaction over
//            original ids; counts stay on the device (the reference copies
//            them to the host every step, cu:2532-2540).
//   seed     initialize_sub_particles (cu:1023-1090) + project_quantities
//            (cu:570-622) for the boundary sub-particles.
//   sort     sub hash + sort + cell ranges on the sub grid (cu:121-153,
//            2327-2389): the counting sort of cell_sort.hip, then each cell
//            ranked back into list order (split parents then boundary
//            parents, by id, children in order: the reference's tie order);
//            the stable radix sort for sub grids too large for it.
//   physics  sub density (split subs), sub acceleration (split subs), sub
//            integration + collision (all subs) (cu:259-313, 477-568,
//            667-696, 836-930); the density pass records each split sub's
//            neighbours as packed row words and the acceleration pass runs
//            its pair terms over them.
// Every kernel reads the live counts from device memory; grids are sized
// for the sub-particle capacity.
#include <algorithm>
#include <cstring>
#include <vector>

#include "engine.h"

namespace sph {

struct TSCounts {
    uint32_t nsplit, nboundary, ndefault, nsub, nparents, nsub_prev;
    uint32_t nmark;   // entries of the split list of this step (k_ts_split or the hit-mask
                      // density pass); zeroed by k_ts_force
    uint32_t pad1;
};

struct TwoScale {
    uint32_t cap = 0;            // sub-particle capacity
    uint32_t nblk = 0;           // compaction blocks
    uint32_t sub_cells = 0;
    int sub_bits = 1;
    uint8_t* state_fin = nullptr;   // final state, sorted order
    uint32_t* mark_list = nullptr;  // [N] slots of this step's split particles, in no order
    uint8_t* cls = nullptr;         // class by original id: 0 default, 1 split, 2 boundary
    uint32_t *split_ids = nullptr, *boundary_ids = nullptr, *default_ids = nullptr;
    uint32_t* blk = nullptr;        // [3][nblk] block counts -> offsets
    TSCounts* counts = nullptr;
    float4 *spos = nullptr, *svel = nullptr;          // list order (seeded)
    uint32_t *sid = nullptr, *skey = nullptr;         // list order
    uint32_t *sskey = nullptr, *sperm = nullptr;      // sorted order
    float4 *spos_s = nullptr, *svel_s = nullptr;      // sorted order
    float4 *spos_out = nullptr, *svel_out = nullptr, *sacc = nullptr;  // sorted, after the step
    int2* srange = nullptr;
    RadixSort sorter;
    // counting sort over the sub grid (used when it has at most
    // kCellSortRatio cells per sub-particle slot); its order inside a cell is
    // replaced by list order in k_ts_gather_ranked
    CellSort csort;
    bool use_cell_sort = false;
    uint32_t* sperm_raw = nullptr;
    uint64_t* smasks = nullptr;   // row words of the split subs (format in common.h; sub density -> sub force)
    size_t bytes = 0;
    bool sub_cells_valid = false;
};

namespace {

constexpr int kB = 256;
constexpr int kCompactItems = 4;
constexpr int kCompactTile = kB * kCompactItems;

inline unsigned nblocks(uint32_t n) { return (n + kB - 1) / kB; }

// shift_state + find_split_particles, sorted order (own slot only); the
// final state starts as the shifted state (the boundary bit is set by
// k_ts_mark), and the split particles' slots are appended to the split list
// (one atomic per wavefront; the list order does not matter to k_ts_mark).
__global__ __launch_bounds__(kB) void k_ts_split(const float4* __restrict__ pos, uint8_t* __restrict__ state,
                                                 uint8_t* __restrict__ state_fin, uint32_t n,
                                                 const DevScene* __restrict__ scenes, TSCounts* __restrict__ cnt,
                                                 uint32_t* __restrict__ mark_list) {
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if ((s & ~63u) >= n) return;   // the whole wavefront lies past the end
    const int lane = threadIdx.x & 63;
    const bool valid = s < n;
    bool in = false;
    if (valid) {
        const DevScene& S = scenes[0];
        float4 x = pos[s];
        uint8_t st = (uint8_t)(state[s] << 2);
        in = x.x >= S.split_lo[0] && x.x <= S.split_hi[0] && x.y >= S.split_lo[1] && x.y <= S.split_hi[1] &&
             x.z >= S.split_lo[2] && x.z <= S.split_hi[2];
        if (in) st |= 1;
        state[s] = st;
        state_fin[s] = st;
    }
    const uint64_t ins = __ballot(in);
    if (!ins) return;
    const int leader = __ffsll((unsigned long long)ins) - 1;
    uint32_t base = 0;
    if (lane == leader) base = atomicAdd(&cnt->nmark, (uint32_t)__popcll(ins));
    base = __shfl(base, leader, 64);
    if (in) {
        const uint64_t below = lane == 0 ? 0ull : (ins & (~0ull >> (64 - lane)));
        mark_list[base + (uint32_t)__popcll(below)] = s;
    }
}

// find_boundary_particles (bit 2: a particle with its split bit set this
// step lies within h) as the reference's push from the split particles,
// which are few: each sets the bit of every particle within h of it, itself
// included.  Every writer of a byte writes the same value (the particle's
// state | 2), so the order of the writes does not matter.
// One wavefront per split particle (a fixed grid of wavefronts strides over
// the split list): lane l takes neighbour row l / 7 (nine rows as a rule;
// rows past 8 in the rare four-cell spans go to lanes l / 7 + 9, ...) and
// every seventh candidate of it from l % 7.  Each lane then has a short
// chain of dependent loads (one range round, ~4 candidates); with one thread
// per split particle the chain ran through all ~216 candidates and the pass
// waited on memory (38 us at 262,144 particles with eight lanes each).
constexpr int kMarkRowLanes = 7;
constexpr unsigned kMarkGrid = 1024;   // blocks: 4,096 wavefronts

__global__ __launch_bounds__(kB) void k_ts_mark(const float4* __restrict__ pos, const uint8_t* __restrict__ state,
                                                const int2* __restrict__ range, const DevScene* __restrict__ scenes,
                                                const TSCounts* __restrict__ cnt,
                                                const uint32_t* __restrict__ mark_list,
                                                uint8_t* __restrict__ state_fin) {
    const uint32_t nmark = cnt->nmark;
    const int lane = threadIdx.x & 63;
    const uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const uint32_t nwaves = (gridDim.x * blockDim.x) >> 6;
    const DevScene& S = scenes[0];
    const float h = S.h, h2 = S.h2;
    for (uint32_t w = wave; w < nmark; w += nwaves) {
        const uint32_t s = mark_list[w];
        if (lane == 0) state_fin[s] = (uint8_t)(state[s] | 2);   // within h of itself
        if (lane >= 9 * kMarkRowLanes) continue;
        const float4 xi = pos[s];
        const int lx = range_coord(xi.x, -h, S.ox, S.cell, S.nx), hx = range_coord(xi.x, h, S.ox, S.cell, S.nx);
        const int ly = range_coord(xi.y, -h, S.oy, S.cell, S.ny), hy = range_coord(xi.y, h, S.oy, S.cell, S.ny);
        const int lz = range_coord(xi.z, -h, S.oz, S.cell, S.nz), hz = range_coord(xi.z, h, S.oz, S.cell, S.nz);
        if (hy < ly || hz < lz) continue;   // a range wrapped by the grid's "% dim": no rows
        const int ny_rows = hy - ly + 1, nrows = (hz - lz + 1) * ny_rows;
        const int sub = lane % kMarkRowLanes;
        for (int row = lane / kMarkRowLanes; row < nrows; row += 9) {
            const int k = lz + row / ny_rows, j = ly + row % ny_rows;
            const uint32_t rowbase = S.cell_offset + (uint32_t)(S.nx * (S.ny * k + j));
            int b = 0x7fffffff, e = -1;
            for (int i = lx; i <= hx; ++i) {
                const int2 r = range[rowbase + (uint32_t)i];
                if (r.x >= 0) {
                    b = min(b, r.x);
                    e = max(e, r.y);
                }
            }
            if (b >= e) continue;   // empty row (b = INT_MAX: b + sub must not be formed)
            for (int q = b + sub; q < e; q += kMarkRowLanes) {
                // split particles mark themselves (above): the stores go to
                // the few non-split particles of the box's shell only
                const uint8_t sq = state[q];
                if (sq & 1) continue;
                const float4 xj = pos[q];
                const float dx = xj.x - xi.x, dy = xj.y - xi.y, dz = xj.z - xi.z;
                if (dx * dx + dy * dy + dz * dz <= h2) state_fin[q] = (uint8_t)(sq | 2);
            }
        }
    }
}

// The same push driven by the coarse density pass's row words (first slot,
// masks of the row's candidates within h in chunks of 32, own slot
// excluded; kRowLong for a row of more than kRowMaxCands candidates), when
// the step runs the hit-mask pair: one lane per split particle and row, no
// cell-table reads or distance tests for all but kRowLong rows.

__global__ __launch_bounds__(kB) void k_ts_mark_rows(const float4* __restrict__ pos, const uint8_t* __restrict__ state,
                                                     const int2* __restrict__ range, const DevScene* __restrict__ scenes,
                                                     const TSCounts* __restrict__ cnt,
                                                     const uint32_t* __restrict__ mark_list,
                                                     const uint64_t* __restrict__ masks, uint32_t n,
                                                     uint8_t* __restrict__ state_fin) {
    const uint64_t total = 9ull * cnt->nmark;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; g < total; g += stride) {
        const uint32_t s = mark_list[g / 9];
        const int row = (int)(g % 9);
        if (row == 0) state_fin[s] = (uint8_t)(state[s] | 2);   // within h of itself
        const uint64_t w = masks[(size_t)row * n + s];
        if (w != kRowLong) {
            const uint32_t last = row_last_chunk(w);
            for (uint32_t c = 0; c <= last; ++c) {
                const uint32_t b = row_first(w) + 32u * c;
                uint32_t m = c == 0 ? (uint32_t)w : chunk_masks(masks, n)[chunk_at(n, c, row, s)];
                while (m) {
                    const uint32_t q = b + (uint32_t)(__ffs(m) - 1);
                    m &= m - 1;
                    const uint8_t sq = state[q];
                    if (!(sq & 1)) state_fin[q] = (uint8_t)(sq | 2);
                }
            }
            continue;
        }
        // long row: its candidates tested here
        const DevScene& S = scenes[0];
        const float h = S.h, h2 = S.h2;
        const float4 xi = pos[s];
        const int lx = range_coord(xi.x, -h, S.ox, S.cell, S.nx), hx = range_coord(xi.x, h, S.ox, S.cell, S.nx);
        const int ly = range_coord(xi.y, -h, S.oy, S.cell, S.ny), hy = range_coord(xi.y, h, S.oy, S.cell, S.ny);
        const int lz = range_coord(xi.z, -h, S.oz, S.cell, S.nz), hz = range_coord(xi.z, h, S.oz, S.cell, S.nz);
        const int k = lz + row / 3, j = ly + row % 3;
        if (k > hz || j > hy) continue;
        const uint32_t rowbase = S.cell_offset + (uint32_t)(S.nx * (S.ny * k + j));
        int b = 0x7fffffff, e = -1;
        for (int i = lx; i <= hx; ++i) {
            const int2 r = range[rowbase + (uint32_t)i];
            if (r.x >= 0) {
                b = min(b, r.x);
                e = max(e, r.y);
            }
        }
        for (int q = b; q < e; ++q) {
            const uint8_t sq = state[q];
            if (sq & 1) continue;
            const float4 xj = pos[q];
            const float dx = xj.x - xi.x, dy = xj.y - xi.y, dz = xj.z - xi.z;
            if (dx * dx + dy * dy + dz * dz <= h2) state_fin[q] = (uint8_t)(sq | 2);
        }
    }
}

// The class of each particle by original id: 0 default, 1 split (split bit,
// boundary bit set by itself), 2 boundary only.
__global__ __launch_bounds__(kB) void k_ts_classify(const uint8_t* __restrict__ state_fin,
                                                    const uint32_t* __restrict__ oid, uint32_t n,
                                                    uint8_t* __restrict__ cls) {
    uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    const int c = state_fin[s] & 3;
    cls[oid[s]] = c == 0 ? 0 : (c == 2 ? 2 : 1);
}

// 3-way compaction in original-id order: per-tile counts ...
__global__ __launch_bounds__(kB) void k_ts_count(const uint8_t* __restrict__ cls, uint32_t n,
                                                 uint32_t* __restrict__ blk, uint32_t nblk) {
    __shared__ uint32_t s_c[3];
    if (threadIdx.x < 3) s_c[threadIdx.x] = 0;
    __syncthreads();
    uint32_t c[3] = {0, 0, 0};
    uint32_t base = blockIdx.x * kCompactTile;
    for (int r = 0; r < kCompactItems; ++r) {
        uint32_t i = base + r * kB + threadIdx.x;
        if (i < n) c[cls[i]]++;
    }
    for (int t = 0; t < 3; ++t) {
        uint32_t v = c[t];
        for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
        if ((threadIdx.x & 63) == 0 && v) atomicAdd(&s_c[t], v);
    }
    __syncthreads();
    if (threadIdx.x < 3) blk[threadIdx.x * nblk + blockIdx.x] = s_c[threadIdx.x];
}

// ... and every element writes its id into its list (collect_ids).  Each
// block sums the tile counts before its own (its list offsets) and all of
// them (the totals and the capacity check) itself, so no separate scan pass
// runs; block 0 publishes the counts.
__global__ __launch_bounds__(kB) void k_ts_collect(const uint8_t* __restrict__ cls, uint32_t n,
                                                   const uint32_t* __restrict__ blk, uint32_t nblk, uint32_t cap,
                                                   TSCounts* __restrict__ cnt, uint32_t* __restrict__ status,
                                                   uint32_t* __restrict__ split_ids,
                                                   uint32_t* __restrict__ boundary_ids,
                                                   uint32_t* __restrict__ default_ids) {
    __shared__ uint32_t s_w[3][kB / 64];
    __shared__ uint32_t s_carry[3];
    __shared__ uint32_t s_red[2][3][kB / 64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    // offsets (tiles before this one) and totals of the three classes
    uint32_t before[3] = {0, 0, 0}, all3[3] = {0, 0, 0};
    for (uint32_t b = threadIdx.x; b < nblk; b += kB)
        for (int t = 0; t < 3; ++t) {
            const uint32_t v = blk[t * nblk + b];
            all3[t] += v;
            if (b < blockIdx.x) before[t] += v;
        }
    for (int t = 0; t < 3; ++t)
        for (int o = 32; o > 0; o >>= 1) {
            before[t] += __shfl_down(before[t], o, 64);
            all3[t] += __shfl_down(all3[t], o, 64);
        }
    if (lane == 0)
        for (int t = 0; t < 3; ++t) {
            s_red[0][t][wave] = before[t];
            s_red[1][t][wave] = all3[t];
        }
    __syncthreads();
    uint32_t tot[3];
    for (int t = 0; t < 3; ++t) {
        uint32_t a = 0, c = 0;
        for (int w = 0; w < kB / 64; ++w) {
            a += s_red[0][t][w];
            c += s_red[1][t][w];
        }
        if (threadIdx.x == 0) s_carry[t] = a;
        tot[t] = c;
    }
    // capacity: keep whole parents that fit, split parents first
    uint32_t nsplit = tot[1], nbound = tot[2];
    uint64_t nsub = 8ull * (nsplit + nbound);
    const bool over = nsub > cap;
    if (over) {
        const uint32_t fit = cap / 8;
        if (nsplit > fit) { nsplit = fit; nbound = 0; }
        else if (nsplit + nbound > fit) nbound = fit - nsplit;
        nsub = 8ull * (nsplit + nbound);
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        if (over) atomicOr(status, kStatusCapacity);
        cnt->nsub_prev = cnt->nsub;
        cnt->nsplit = nsplit;
        cnt->nboundary = nbound;
        cnt->ndefault = tot[0];
        cnt->nparents = nsplit + nbound;
        cnt->nsub = (uint32_t)nsub;
    }
    __syncthreads();
    uint32_t base = blockIdx.x * kCompactTile;
    for (int r = 0; r < kCompactItems; ++r) {
        uint32_t i = base + r * kB + threadIdx.x;
        int c = i < n ? cls[i] : -1;
        uint32_t rank[3], tt[3];
        for (int t = 0; t < 3; ++t) {
            uint64_t m = __ballot(c == t);
            uint64_t below = lane == 0 ? 0ull : (~0ull >> (64 - lane));
            rank[t] = (uint32_t)__popcll(m & below);
            if (lane == 0) s_w[t][wave] = (uint32_t)__popcll(m);
        }
        __syncthreads();
        for (int t = 0; t < 3; ++t) {
            uint32_t off = 0, all = 0;
            for (int w = 0; w < kB / 64; ++w) {
                if (w < wave) off += s_w[t][w];
                all += s_w[t][w];
            }
            rank[t] += off + s_carry[t];
            tt[t] = all;
        }
        if (c == 1 && rank[1] < nsplit) split_ids[rank[1]] = i;
        if (c == 2 && rank[2] < nbound) boundary_ids[rank[2]] = i;
        if (c == 0) default_ids[rank[0]] = i;
        __syncthreads();
        if (threadIdx.x < 3) s_carry[threadIdx.x] += tt[threadIdx.x];
        __syncthreads();
    }
}

// initialize_sub_particles for every split and boundary parent (cu:1023-
// 1090): 8 children, directions (+-1,+-1,+-1) in the order of the reference
// table (x sign from bit 2, y from bit 1, z from bit 0 of the child index).
// The sub cell keys of the seeded positions (compute_sub_particle_hash,
// cu:121-153) are computed here too, and (clear_prev) the cells of the
// previous step's sorted sub keys are invalidated before the sub sort
// writes this step's ranges (what k_ts_hash does for a sub-particle step).
__global__ __launch_bounds__(kB) void k_ts_seed(const TSCounts* __restrict__ cnt,
                                                const uint32_t* __restrict__ split_ids,
                                                const uint32_t* __restrict__ boundary_ids,
                                                const uint32_t* __restrict__ slot_of_oid,
                                                const float4* __restrict__ pos, const float4* __restrict__ vel,
                                                const DevScene* __restrict__ scenes, float4* __restrict__ spos,
                                                float4* __restrict__ svel, uint32_t* __restrict__ sid,
                                                uint32_t* __restrict__ skey, uint32_t* __restrict__ status,
                                                int clear_prev, const uint32_t* __restrict__ sskey,
                                                int2* __restrict__ srange) {
    uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (clear_prev && k < cnt->nsub_prev) {
        const uint32_t pc = sskey[k];
        if (k == 0 || sskey[k - 1] != pc) srange[pc] = make_int2(-1, -1);
    }
    const uint32_t nsub = cnt->nsub, nsplit = cnt->nsplit;
    if (k >= nsub) return;
    const DevScene& S = scenes[0];
    uint32_t q = k >> 3, c = k & 7;
    uint32_t id = q < nsplit ? split_ids[q] : boundary_ids[q - nsplit];
    uint32_t slot = slot_of_oid[id];
    float4 xp = pos[slot];
    float4 vp = vel[slot];
    float radius, dirf;
    if (S.sub_placement == 0) {
        // r = (3 m / (4 pi rho))^(1/3), along the cube diagonals
        float radicand = 3.0f * S.mass / (4.0f * 3.14159265358979323846f * vp.w);
        radius = cbrtf(radicand);
        dirf = 0.577350269f;   // 1/sqrt(3) (cu:1025)
    } else {
        radius = 0.25f * S.h;
        dirf = 1.0f;
    }
    float off = radius * dirf;
    float ox = (c & 4) ? -off : off, oy = (c & 2) ? -off : off, oz = (c & 1) ? -off : off;
    const float4 xs = make_float4(xp.x + ox, xp.y + oy, xp.z + oz, 0.0f);
    spos[k] = xs;
    svel[k] = make_float4(vp.x, vp.y, vp.z, 0.0f);
    sid[k] = 8u * id + c;
    uint32_t flags = 0;
    skey[k] = sub_cell_key(S, xs.x, xs.y, xs.z, flags);
    if (flags) atomicOr(status, flags);
}

// project_quantities (cu:570-622, 1328-1381): boundary sub-particles take the
// unweighted mean density of the coarse particles within h_sub, searched on
// the coarse grid; the mean pressure is k (mean density - rho0) since the
// equation of state is linear.  kProjLanes lanes per boundary sub-particle,
// one coarse cell each (the 2x2x2 cells a sub-particle's range spans as a
// rule; lanes loop over larger spans), over a fixed grid striding over the
// boundary subs: each lane's chain of dependent loads is one cell long
// instead of eight.  The group adds sums and counts by shuffles.
constexpr int kProjLanes = 8;
constexpr unsigned kProjGrid = 512;

__device__ __forceinline__ float proj_sum(float v) {
#pragma unroll
    for (int m = 1; m < kProjLanes; m <<= 1) v += __shfl_xor(v, m, 64);
    return v;
}

__global__ __launch_bounds__(kB) void k_ts_project(const TSCounts* __restrict__ cnt, const float4* __restrict__ spos,
                                                   float4* __restrict__ svel, const float4* __restrict__ pos,
                                                   const float4* __restrict__ vel, const int2* __restrict__ range,
                                                   const DevScene* __restrict__ scenes) {
    const uint32_t nsub = cnt->nsub, first = 8u * cnt->nsplit;
    const uint64_t total = (uint64_t)(nsub > first ? nsub - first : 0u) * kProjLanes;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;   // a multiple of kProjLanes
    const DevScene& S = scenes[0];
    const float hs = S.h_sub, hs2 = S.h2_sub;
    for (uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; g < total; g += stride) {
        const uint32_t k = first + (uint32_t)(g / kProjLanes);
        const int lane = (int)(g % kProjLanes);
        const float4 xs = spos[k];
        const int lx = range_coord(xs.x, -hs, S.ox, S.cell, S.nx), hx = range_coord(xs.x, hs, S.ox, S.cell, S.nx);
        const int ly = range_coord(xs.y, -hs, S.oy, S.cell, S.ny), hy = range_coord(xs.y, hs, S.oy, S.cell, S.ny);
        const int lz = range_coord(xs.z, -hs, S.oz, S.cell, S.nz), hz = range_coord(xs.z, hs, S.oz, S.cell, S.nz);
        float sum = 0.0f, cntn = 0.0f;
        if (hx >= lx && hy >= ly && hz >= lz) {   // else a range wrapped by the grid's "% dim": no cells
            const int nxc = hx - lx + 1, nyc = hy - ly + 1, ncell = nxc * nyc * (hz - lz + 1);
            for (int c = lane; c < ncell; c += kProjLanes) {
                const int i = lx + c % nxc, j = ly + (c / nxc) % nyc, kk = lz + c / (nxc * nyc);
                const int2 r = range[S.cell_offset + (uint32_t)(S.nx * (S.ny * kk + j) + i)];
                for (int q = r.x; q < r.y; ++q) {
                    const float4 xj = pos[q];
                    const float dx = xs.x - xj.x, dy = xs.y - xj.y, dz = xs.z - xj.z;
                    if (dx * dx + dy * dy + dz * dz <= hs2) {
                        sum += vel[q].w;
                        cntn += 1.0f;
                    }
                }
            }
        }
        sum = proj_sum(sum);     // every lane of the group takes part
        cntn = proj_sum(cntn);
        if (lane == 0) svel[k].w = cntn > 0.0f ? sum / cntn : S.rho0;
    }
}

// Sub cell keys of this step's sub-particles, and (clear_prev) the
// invalidation of the cells of the previous step's sorted sub keys before
// the sub sort writes this step's ranges.
__global__ __launch_bounds__(kB) void k_ts_hash(const TSCounts* __restrict__ cnt, const float4* __restrict__ spos,
                                                uint32_t* __restrict__ skey, const DevScene* __restrict__ scenes,
                                                uint32_t* __restrict__ status, int clear_prev,
                                                const uint32_t* __restrict__ sskey, int2* __restrict__ srange) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (clear_prev && k < cnt->nsub_prev) {
        const uint32_t c = sskey[k];
        if (k == 0 || sskey[k - 1] != c) srange[c] = make_int2(-1, -1);
    }
    if (k >= cnt->nsub) return;
    float4 x = spos[k];
    uint32_t flags = 0;
    skey[k] = sub_cell_key(scenes[0], x.x, x.y, x.z, flags);
    if (flags) atomicOr(status, flags);
}

// sub cell ranges + gather into sub-sorted order (stable list order = the
// reference's tie order, no fix-up needed).
__global__ __launch_bounds__(kB) void k_ts_cells_gather(const TSCounts* __restrict__ cnt, const uint32_t* __restrict__ sskey,
                                                        const uint32_t* __restrict__ sperm,
                                                        const float4* __restrict__ spos, const float4* __restrict__ svel,
                                                        int2* __restrict__ srange, float4* __restrict__ spos_s,
                                                        float4* __restrict__ svel_s) {
    uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t n = cnt->nsub;
    if (t >= n) return;
    uint32_t k = sskey[t];
    int* r = reinterpret_cast<int*>(srange);
    if (t == 0 || sskey[t - 1] != k) r[2 * (size_t)k] = (int)t;
    if (t == n - 1 || sskey[t + 1] != k) r[2 * (size_t)k + 1] = (int)(t + 1);
    uint32_t src = sperm[t];
    spos_s[t] = spos[src];
    svel_s[t] = svel[src];
}

// After the counting sort: each sub cell's slots in list order (the
// reference's stable tie order: the rank of the list index among the cell's
// slots), then the gather into sorted order.  The cell ranges come from the
// sort.
__global__ __launch_bounds__(kB) void k_ts_gather_ranked(const TSCounts* __restrict__ cnt,
                                                         const uint32_t* __restrict__ sskey,
                                                         const uint32_t* __restrict__ sperm_raw,
                                                         const int2* __restrict__ srange,
                                                         const float4* __restrict__ spos,
                                                         const float4* __restrict__ svel, uint32_t* __restrict__ sperm,
                                                         float4* __restrict__ spos_s, float4* __restrict__ svel_s) {
    uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= cnt->nsub) return;
    const int2 r = srange[sskey[t]];
    const uint32_t src = sperm_raw[t];
    uint32_t rank = 0;
    for (int q = r.x; q < r.y; ++q) rank += sperm_raw[q] < src ? 1u : 0u;
    const uint32_t dst = (uint32_t)r.x + rank;
    sperm[dst] = src;
    spos_s[dst] = spos[src];
    svel_s[dst] = svel[src];
}

__device__ __forceinline__ void sub_range(const DevScene& S, float4 x, float h, int& lx, int& hx, int& ly, int& hy,
                                          int& lz, int& hz) {
    lx = range_coord(x.x, -h, S.ox, S.cell_sub, S.snx); hx = range_coord(x.x, h, S.ox, S.cell_sub, S.snx);
    ly = range_coord(x.y, -h, S.oy, S.cell_sub, S.sny); hy = range_coord(x.y, h, S.oy, S.cell_sub, S.sny);
    lz = range_coord(x.z, -h, S.oz, S.cell_sub, S.snz); hz = range_coord(x.z, h, S.oz, S.cell_sub, S.snz);
}

// The sub density pass gives each sub-particle kTsLanes consecutive
// lanes: lane j of a group takes candidates b + j, b + j + 4, ...
// of every row run and the group adds its partial sums by shuffles.  The
// sub population (~10^5) is too small to hide memory latency at one thread
// per particle.  A fixed grid strides over nsub * kTsLanes lanes (the count
// is read on the device); the stride is a multiple of kTsLanes, so the lanes
// of a group stay together.  (Eight lanes per sub-particle measured 49 us
// against 35; per-wavefront LDS windows of the candidates, as in the coarse
// density pass, 44 us: a wavefront here holds only 16 sub-particles.)
constexpr int kTsLanes = 4;
constexpr unsigned kTsGrid = 2048;

__device__ __forceinline__ float group_sum(float v) {
#pragma unroll
    for (int m = 1; m < kTsLanes; m <<= 1) v += __shfl_xor(v, m, 64);
    return v;
}

// The nine x-row runs of a sub-particle's neighbourhood on the sub grid: the
// sub grid's cells (size cell_sub), support h_sub, the same 3 x 3 row set
// and y-z cull as the coarse passes.
struct SubRows {
    int lx, hx, ly, hy, lz, hz;
};

__device__ __forceinline__ SubRows sub_rows(const DevScene& S, float4 x) {
    SubRows r;
    sub_range(S, x, S.h_sub, r.lx, r.hx, r.ly, r.hy, r.lz, r.hz);
    return r;
}

__device__ __forceinline__ void sub_row_slots(const int2* __restrict__ srange, const DevScene& S, const SubRows& R,
                                              int dk, int dj, float4 xi, int& b, int& e) {
    const int k = R.lz + dk, j = R.ly + dj;
    b = 0;
    e = 0;
    if (k > R.hz || j > R.hy) return;
    const float y0 = S.oy + (float)j * S.cell_sub, z0 = S.oz + (float)k * S.cell_sub;
    const float dy = fmaxf(fmaxf(y0 - xi.y, xi.y - (y0 + S.cell_sub)), 0.0f);
    const float dz = fmaxf(fmaxf(z0 - xi.z, xi.z - (z0 + S.cell_sub)), 0.0f);
    if ((dy * dy + dz * dz) * 0.99999f > S.h2_sub) return;
    const uint32_t rowbase = (uint32_t)(S.snx * (S.sny * k + j));
    b = 0x7fffffff;
    e = -1;
    for (int i = R.lx; i <= R.hx; ++i) {
        const int2 r = srange[rowbase + (uint32_t)i];
        if (r.x >= 0) {
            b = min(b, r.x);
            e = max(e, r.y);
        }
    }
    if (e < b) { b = 0; e = 0; }
}

// Row words of the sub passes: the coarse pair's format (common.h) over the
// sub grid -- first slot, the force-neighbour masks of up to four chunks of
// 32 candidates, kRowLong for a row of more than kRowMaxCands.

// compute_sub_particle_density_pressure (cu:259-313): split subs only, no
// wall term.  Besides the density, each split sub's group records per row the
// candidates within h_sub (r^2 <= h_sub^2, its own slot excluded) as a
// packed row word, so that the force pass runs its pair terms over those
// bits without distance tests or cell-table reads.
__global__ __launch_bounds__(kB) void k_ts_density(const TSCounts* __restrict__ cnt, const uint32_t* __restrict__ sperm,
                                                   const float4* __restrict__ spos_s, float4* __restrict__ svel_s,
                                                   const int2* __restrict__ srange,
                                                   const DevScene* __restrict__ scenes, uint32_t cap,
                                                   uint64_t* __restrict__ smasks) {
    const uint64_t total = (uint64_t)cnt->nsub * kTsLanes;
    const uint32_t nsplit8 = 8u * cnt->nsplit;
    const DevScene& S = scenes[0];
    const float h2 = S.h2_sub;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; g < total; g += stride) {
        const uint32_t t = (uint32_t)(g / kTsLanes);
        const int lane = (int)(g % kTsLanes);
        const bool split = sperm[t] < nsplit8;   // boundary subs: projected density
        float sum = 0.0f;
        const float4 xi = spos_s[t];
        const F4Buf sp(spos_s);
        const SubRows R = sub_rows(S, xi);
#pragma unroll 1
        for (int row = 0; row < 9; ++row) {
            int b = 0, e = 0;
            if (split) sub_row_slots(srange, S, R, row / 3, row % 3, xi, b, e);
            // the sign bits of h_sub^2 - r^2 of this lane's candidates
            // b + lane + 4k, newest in bit 0 (as in the coarse density pass)
            uint32_t far = 0;
#pragma unroll 4
            for (int q = b + lane; q < e; q += kTsLanes) {
                const float4 xj = sp[q];
                const float dx = xj.x - xi.x, dy = xj.y - xi.y, dz = xj.z - xi.z;
                float d = fmaf(-dx, dx, fmaf(-dy, dy, fmaf(-dz, dz, h2)));
                far = __builtin_amdgcn_alignbit(far, __float_as_uint(d), 31);
                d = fmaxf(d, 0.0f);
                sum += d * d * d;
            }
            // a row of at most 128 candidates gives each lane at most 32: bit
            // k of the lane's near bits goes to candidate position 4k + lane,
            // chunk k / 8 of the row (the group's lanes share the row, so
            // they run the chunk loop and its shuffles together)
            const int cnt = e - b;
            const int c = cnt > lane ? (cnt - lane + kTsLanes - 1) / kTsLanes : 0;
            const uint32_t near = c > 0 && cnt <= kRowMaxCands ? __builtin_bitreverse32(~far) >> (32 - c) : 0u;
            const uint32_t self = t - (uint32_t)b;   // the sub's own slot is no neighbour
            const int nch = cnt <= kRowMaxCands ? max((cnt + 31) >> 5, 1) : 1;
            uint32_t m0 = 0;
            for (int ch = 0; ch < nch; ++ch) {
                uint32_t x = (near >> (8 * ch)) & 0xFFu;
                x = (x | (x << 12)) & 0x000F000Fu;
                x = (x | (x << 6)) & 0x03030303u;
                x = (x | (x << 3)) & 0x11111111u;
                uint32_t m = x << lane;
                m |= __shfl_xor(m, 1, 64);
                m |= __shfl_xor(m, 2, 64);
                if ((self >> 5) == (uint32_t)ch) m &= ~(1u << (self & 31u));
                if (ch == 0) m0 = m;
                else if (split && lane == 0) chunk_masks(smasks, cap)[chunk_at(cap, (uint32_t)ch, row, t)] = m;
            }
            if (split && lane == 0)
                smasks[(size_t)row * cap + t] =
                    cnt <= kRowMaxCands ? ((uint64_t)(uint32_t)b << 32 | m0 | (uint64_t)(nch - 1) << 60) : kRowLong;
        }
        sum = group_sum(sum);   // every lane of the group takes part
        if (split && lane == 0) reinterpret_cast<float*>(svel_s)[4 * (size_t)t + 3] = sum * (S.poly6_sub * S.mass_sub);
    }
}

// compute_sub_particle_acceleration (cu:477-568) for split subs (boundary
// subs: zero acceleration), then integrate_sub_particles_euler (cu:667-696)
// and the collision response for every sub-particle.  One thread per
// sub-particle.  The pair terms run over the set bits of the density pass's
// row words, in candidate order (rows in z-major, y-minor order, slots
// ascending), so a wavefront executes them as often as its lane with the most
// neighbours in the row; a row of more than kRowMaxCands candidates is tested
// here.  (Before the rows kept chunks past 32 candidates, such rows were
// tested here too: on the two-scale workload after ~1,000 steps, as the
// flow compresses, this pass took 120 us against 44 us now.)
__global__ __launch_bounds__(kB) void k_ts_force(TSCounts* __restrict__ cnt, const uint32_t* __restrict__ sperm,
                                                 const float4* __restrict__ spos_s, const float4* __restrict__ svel_s,
                                                 const int2* __restrict__ srange, const DevScene* __restrict__ scenes,
                                                 const uint64_t* __restrict__ smasks, uint32_t cap,
                                                 const float* __restrict__ gridd, float4* __restrict__ spos_out,
                                                 float4* __restrict__ svel_out, float4* __restrict__ sacc,
                                                 uint32_t* __restrict__ status) {
    uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t nsub = cnt->nsub;
    if (t == 0) {
        // the last sub kernel of the step: this step's count becomes the
        // previous one (its cells are cleared next step) and the next
        // step's split list starts empty (no other thread reads either)
        cnt->nsub_prev = nsub;
        cnt->nmark = 0;
    }
    if (t >= nsub) return;
    const DevScene& S = scenes[0];
    const float4 xi = spos_s[t];
    const float4 vi = svel_s[t];
    float ax = 0.0f, ay = 0.0f, az = 0.0f;
    if (sperm[t] < 8u * cnt->nsplit) {
        const float h = S.h_sub, h2 = S.h2_sub;
        const float rhoi = vi.w;
        const float Ai = S.stiff * (rhoi - S.rho0) / rhoi;
        const float kr0 = S.stiff * S.rho0;
        float fpx = 0, fpy = 0, fpz = 0, fvx = 0, fvy = 0, fvz = 0, cgx = 0, cgy = 0, cgz = 0, cl = 0;
        const F4Buf sp(spos_s), sv(svel_s);
        auto pair = [&](int q) {
            float4 xj = sp[q];
            float dx = xi.x - xj.x, dy = xi.y - xj.y, dz = xi.z - xj.z;
            float r2 = dx * dx + dy * dy + dz * dz;
            float4 vj = sv[q];
            // a distinct sub at distance 0 contributes nothing (every term
            // carries 1/r or 1/rho_j, both taken as 0 there)
            const bool nz = r2 > 0.0f;
            float inv_r = nz ? __builtin_amdgcn_rsqf(r2) : 0.0f;
            float rn = r2 * inv_r;
            float invj = nz ? __builtin_amdgcn_rcpf(vj.w) : 0.0f;
            float pj = S.stiff * vj.w - kr0;
            float hr = h - rn, qd = h2 - r2;
            float cp = (Ai + rhoi * pj * invj * invj) * hr * hr * inv_r;
            fpx += cp * dx; fpy += cp * dy; fpz += cp * dz;
            float wv = invj * hr;
            fvx += wv * (vj.x - vi.x); fvy += wv * (vj.y - vi.y); fvz += wv * (vj.z - vi.z);
            float wc = invj * qd * qd;
            cgx += wc * dx; cgy += wc * dy; cgz += wc * dz;
            cl += invj * qd * (3.0f * h2 - 7.0f * r2);
        };
        auto near = [&](int q) {
            const float4 xj = spos_s[q];
            const float dx = xi.x - xj.x, dy = xi.y - xj.y, dz = xi.z - xj.z;
            const float r2 = dx * dx + dy * dy + dz * dz;
            return r2 <= h2 && r2 > 0.0f;
        };
#pragma unroll 1
        for (int row = 0; row < 9; ++row) {
            const uint64_t w = smasks[(size_t)row * cap + t];
            if (w != kRowLong) {
                const uint32_t last = row_last_chunk(w);
                for (uint32_t ch = 0; ch <= last; ++ch) {
                    const int b = (int)(row_first(w) + 32u * ch);
                    uint32_t m = ch == 0 ? (uint32_t)w : chunk_masks(smasks, cap)[chunk_at(cap, ch, row, t)];
                    while (m) {
                        const int u = __ffs(m) - 1;
                        m &= m - 1;
                        pair(b + u);
                    }
                }
            } else {
                int b, e;
                sub_row_slots(srange, S, sub_rows(S, xi), row / 3, row % 3, xi, b, e);
                for (int q = b; q < e; ++q)
                    if (near(q)) pair(q);
            }
        }
        const float cpf = -S.mass_sub * S.grad_spiky_sub, cvf = S.mu * S.mass_sub * S.lapl_visc_sub;
        float Fx = cvf * fvx + cpf * fpx, Fy = cvf * fvy + cpf * fpy, Fz = cvf * fvz + cpf * fpz;
        float cm = S.mass_sub * S.grad_poly6_sub;
        float gx = cm * cgx, gy = cm * cgy, gz = cm * cgz;
        float gnorm = sqrtf(gx * gx + gy * gy + gz * gz);
        if (gnorm > S.norm_thresh) {
            float fc = S.sigma * (S.mass_sub * S.lapl_poly6_sub * cl) / gnorm;
            Fx -= fc * gx; Fy -= fc * gy; Fz -= fc * gz;
        }
        ax = Fx / rhoi;
        ay = Fy / rhoi - S.gravity;
        az = Fz / rhoi;
    }
    sacc[t] = make_float4(ax, ay, az, 0.0f);
    const float dt = S.dt_sub;
    float vx = vi.x + dt * ax, vy = vi.y + dt * ay, vz = vi.z + dt * az;
    float px = xi.x + dt * vx, py = xi.y + dt * vy, pz = xi.z + dt * vz;
    if (S.collision) {
        float cx, cy, cz;
        texel_coords(S, px, py, pz, cx, cy, cz);
        float D = dist_at(gridd, S, cx, cy, cz);
        if (D > 0.0f) {
            float3 gw = dist_grad_normalized(gridd, S, cx, cy, cz);
            gw.x *= S.bscale_x * (1.0f / (float)S.bnx);
            gw.y *= S.bscale_y * (1.0f / (float)S.bny);
            gw.z *= S.bscale_z * (1.0f / (float)S.bnz);
            float gl = sqrtf(gw.x * gw.x + gw.y * gw.y + gw.z * gw.z);
            if (gl > 0.0f) {
                float nx_ = gw.x / gl, ny_ = gw.y / gl, nz_ = gw.z / gl;
                float vn = sqrtf(vx * vx + vy * vy + vz * vz);
                float dp = nx_ * vx + ny_ * vy + nz_ * vz;
                float c = 1.0f;
                if (S.restitution != 0.0f && vn > 0.0f) c += S.restitution * D / (dt * vn);
                c *= dp;
                vx -= c * nx_; vy -= c * ny_; vz -= c * nz_;
                px -= D * nx_; py -= D * ny_; pz -= D * nz_;
            }
        }
    }
    if (!(isfinite(px) && isfinite(py) && isfinite(pz) && isfinite(vx) && isfinite(vy) && isfinite(vz)))
        atomicOr(status, kStatusNonFinite);
    spos_out[t] = make_float4(px, py, pz, 0.0f);
    svel_out[t] = make_float4(vx, vy, vz, vi.w);
}

// The integrated sub-particle state of the last step (sorted order) back to
// list order, as the input of a sub-only step.
__global__ __launch_bounds__(kB) void k_ts_unsort(const TSCounts* __restrict__ cnt, const uint32_t* __restrict__ sperm,
                                                  const float4* __restrict__ spos_out,
                                                  const float4* __restrict__ svel_out, float4* __restrict__ spos,
                                                  float4* __restrict__ svel) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= cnt->nsub) return;
    const uint32_t k = sperm[t];
    spos[k] = spos_out[t];
    svel[k] = svel_out[t];
}

// Sub-particle positions of the last step by sub id, for a renderer:
// (x, y, z, density) of sorted slot q lands at out[sid[sperm[q]]].
__global__ __launch_bounds__(kB) void k_ts_export(const TSCounts* __restrict__ cnt, const float4* __restrict__ spos,
                                                  const float4* __restrict__ svel,
                                                  const uint32_t* __restrict__ sperm,
                                                  const uint32_t* __restrict__ sid, float4* __restrict__ out) {
    uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= cnt->nsub) return;
    const float4 p = spos[q];
    out[sid[sperm[q]]] = make_float4(p.x, p.y, p.z, svel[q].w);
}

}  // namespace

template <class T>
static hipError_t ts_alloc(TwoScale& t, T** p, size_t n) {
    size_t b = sizeof(T) * (n ? n : 1);
    t.bytes += b;
    return hipMalloc(reinterpret_cast<void**>(p), b);
}

sph_status twoscale_create(Engine& e) {
    if (e.nscenes != 1) return e.fail(SPH_INVALID_ARG, "two-scale pass runs on single-scene handles");
    const sph_params& P = e.params[0];
    for (int a = 0; a < 3; ++a)
        if (P.grid_dim_sub[a] < 1) return e.fail(SPH_INVALID_ARG, "sub grid dims must be >= 1");
    if (!(P.grid_spacing_sub > 0.0f) || !(P.compact_support_sub > 0.0f) || !(P.sub_particle_mass > 0.0f))
        return e.fail(SPH_INVALID_ARG, "sub-particle parameters must be positive");
    double sc = (double)P.grid_dim_sub[0] * P.grid_dim_sub[1] * P.grid_dim_sub[2];
    if (sc > 2147483647.0) return e.fail(SPH_INVALID_ARG, "sub grid too large");
    TwoScale* t = new TwoScale();
    uint64_t cap = P.max_sub_particles ? P.max_sub_particles : 8ull * e.N;
    cap = (cap / 8) * 8;
    if (cap < 8) cap = 8;
    if (cap > (1ull << 28) - 8) cap = (1ull << 28) - 8;   // the sub gathers use 32-bit byte offsets
    t->cap = (uint32_t)cap;
    t->nblk = (e.N + kCompactTile - 1) / kCompactTile;
    t->sub_cells = (uint32_t)sc;
    t->sub_bits = 1;
    while (t->sub_bits < 32 && (1ull << t->sub_bits) < (uint64_t)t->sub_cells) ++t->sub_bits;
    bool ok = ts_alloc(*t, &t->state_fin, e.N) == hipSuccess && ts_alloc(*t, &t->cls, e.N) == hipSuccess &&
              ts_alloc(*t, &t->mark_list, e.N) == hipSuccess &&
              ts_alloc(*t, &t->split_ids, e.N) == hipSuccess && ts_alloc(*t, &t->boundary_ids, e.N) == hipSuccess &&
              ts_alloc(*t, &t->default_ids, e.N) == hipSuccess &&
              ts_alloc(*t, &t->blk, 3 * (size_t)t->nblk) == hipSuccess && ts_alloc(*t, &t->counts, 1) == hipSuccess &&
              ts_alloc(*t, &t->spos, t->cap) == hipSuccess && ts_alloc(*t, &t->svel, t->cap) == hipSuccess &&
              ts_alloc(*t, &t->sid, t->cap) == hipSuccess && ts_alloc(*t, &t->skey, t->cap) == hipSuccess &&
              ts_alloc(*t, &t->sskey, t->cap) == hipSuccess && ts_alloc(*t, &t->sperm, t->cap) == hipSuccess &&
              ts_alloc(*t, &t->spos_s, t->cap) == hipSuccess && ts_alloc(*t, &t->svel_s, t->cap) == hipSuccess &&
              ts_alloc(*t, &t->spos_out, t->cap) == hipSuccess && ts_alloc(*t, &t->svel_out, t->cap) == hipSuccess &&
              ts_alloc(*t, &t->sacc, t->cap) == hipSuccess && ts_alloc(*t, &t->srange, t->sub_cells) == hipSuccess &&
              ts_alloc(*t, &t->smasks, row_words_u64(t->cap)) == hipSuccess;
    t->use_cell_sort = (uint64_t)t->sub_cells <= (uint64_t)kCellSortRatio * t->cap;
    if (ok && t->use_cell_sort)
        ok = ts_alloc(*t, &t->sperm_raw, t->cap) == hipSuccess && t->csort.init(t->sub_cells, t->cap) == hipSuccess;
    else if (ok)
        ok = t->sorter.init(t->cap) == hipSuccess;
    e.ts = t;   // owned by the engine from here on (freed by twoscale_destroy)
    if (!ok) {
        (void)hipGetLastError();
        return e.fail(SPH_OUT_OF_MEMORY, "two-scale buffers");
    }
    t->bytes += t->use_cell_sort ? t->csort.bytes : t->sorter.bytes();
    if (hipMemset(t->srange, 0xFF, (size_t)t->sub_cells * sizeof(int2)) != hipSuccess ||
        hipMemset(t->counts, 0, sizeof(TSCounts)) != hipSuccess ||
        hipMemset(t->state_fin, 0, e.N) != hipSuccess || hipDeviceSynchronize() != hipSuccess)
        return e.fail(SPH_DEVICE_ERROR, "two-scale initial clear");
    return SPH_OK;
}

void twoscale_destroy(Engine& e) {
    TwoScale* t = e.ts;
    if (!t) return;
    void* ptrs[] = {t->state_fin, t->cls,   t->split_ids, t->boundary_ids, t->default_ids, t->blk,
                    t->counts,    t->spos,  t->svel,      t->sid,          t->skey,        t->sskey,
                    t->sperm,     t->spos_s, t->svel_s,   t->spos_out,     t->svel_out,    t->sacc,
                    t->srange,    t->sperm_raw, t->mark_list, t->smasks};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    t->sorter.release();
    t->csort.release();
    delete t;
    e.ts = nullptr;
}

// Sub grid and sub physics of the current list-order sub-particles (spos,
// svel): keys, sort, sorted gather, density, acceleration + integration.
static void sub_pass(Engine& e, TwoScale& t, bool keys_ready) {
    hipStream_t st = e.stream;
    if (!keys_ready)
        hipLaunchKernelGGL(k_ts_hash, dim3(nblocks(t.cap)), dim3(kB), 0, st, t.counts, t.spos, t.skey, e.d_scenes,
                           e.status, t.sub_cells_valid ? 1 : 0, t.sskey, t.srange);
    if (t.use_cell_sort) {
        t.csort.sort(st, t.skey, &t.counts->nsub, t.cap, t.sskey, t.sperm_raw, t.srange);
        hipLaunchKernelGGL(k_ts_gather_ranked, dim3(nblocks(t.cap)), dim3(kB), 0, st, t.counts, t.sskey, t.sperm_raw,
                           t.srange, t.spos, t.svel, t.sperm, t.spos_s, t.svel_s);
    } else {
        t.sorter.sort(st, t.skey, nullptr, t.sskey, t.sperm, &t.counts->nsub, t.sub_bits);
        hipLaunchKernelGGL(k_ts_cells_gather, dim3(nblocks(t.cap)), dim3(kB), 0, st, t.counts, t.sskey, t.sperm,
                           t.spos, t.svel, t.srange, t.spos_s, t.svel_s);
    }
    t.sub_cells_valid = true;
    e.mark(kStageTwoScaleSubSort);
    const unsigned sub_grid = (unsigned)std::min<uint64_t>(kTsGrid, ((uint64_t)t.cap * kTsLanes + kB - 1) / kB);
    hipLaunchKernelGGL(k_ts_density, dim3(sub_grid), dim3(kB), 0, st, t.counts, t.sperm, t.spos_s, t.svel_s,
                       t.srange, e.d_scenes, t.cap, t.smasks);
    hipLaunchKernelGGL(k_ts_force, dim3(nblocks(t.cap)), dim3(kB), 0, st, t.counts, t.sperm, t.spos_s, t.svel_s,
                       t.srange, e.d_scenes, t.smasks, t.cap, e.gridd, t.spos_out, t.svel_out, t.sacc, e.status);
    e.mark(kStageTwoScaleSubPhysics);
}

SplitRefs twoscale_split_refs(Engine& e) {
    TwoScale& t = *e.ts;
    return SplitRefs{e.state_b, t.state_fin, &t.counts->nmark, t.mark_list};
}

void twoscale_enqueue(Engine& e, const uint64_t* masks) {
    TwoScale& t = *e.ts;
    hipStream_t st = e.stream;
    const uint32_t n = e.N;
    // with the hit-mask pair the density pass has set the split flags
    if (!masks)
        hipLaunchKernelGGL(k_ts_split, dim3(nblocks(n)), dim3(kB), 0, st, e.pos_b, e.state_b, t.state_fin, n,
                           e.d_scenes, t.counts, t.mark_list);
    if (masks)
        hipLaunchKernelGGL(k_ts_mark_rows, dim3(kMarkGrid), dim3(kB), 0, st, e.pos_b, e.state_b, e.range,
                           e.d_scenes, t.counts, t.mark_list, masks, n, t.state_fin);
    else
        hipLaunchKernelGGL(k_ts_mark, dim3(kMarkGrid), dim3(kB), 0, st, e.pos_b, e.state_b, e.range, e.d_scenes,
                           t.counts, t.mark_list, t.state_fin);
    hipLaunchKernelGGL(k_ts_classify, dim3(nblocks(n)), dim3(kB), 0, st, t.state_fin, e.oid_b, n, t.cls);
    e.mark(kStageTwoScaleState);
    hipLaunchKernelGGL(k_ts_count, dim3(t.nblk), dim3(kB), 0, st, t.cls, n, t.blk, t.nblk);
    hipLaunchKernelGGL(k_ts_collect, dim3(t.nblk), dim3(kB), 0, st, t.cls, n, t.blk, t.nblk, t.cap, t.counts,
                       e.status, t.split_ids, t.boundary_ids, t.default_ids);
    e.mark(kStageTwoScaleCollect);
    hipLaunchKernelGGL(k_ts_seed, dim3(nblocks(t.cap)), dim3(kB), 0, st, t.counts, t.split_ids, t.boundary_ids,
                       e.slot_of_oid, e.pos_b, e.vel_b, e.d_scenes, t.spos, t.svel, t.sid, t.skey, e.status,
                       t.sub_cells_valid ? 1 : 0, t.sskey, t.srange);
    hipLaunchKernelGGL(k_ts_project, dim3(kProjGrid), dim3(kB), 0, st, t.counts, t.spos, t.svel, e.pos_b,
                       e.vel_b, e.range, e.d_scenes);
    e.mark(kStageTwoScaleSeed);
    sub_pass(e, t, true);
}

// AdvanceSubParticles (cu:2119-2143): the sub-particles of the last
// two-scale step advance alone, without reseeding.  Projection takes the
// coarse state as the last step left it, the integrated positions with the
// last sort's cell ranges (the same slots) and the last density pass's
// densities (kept in the positions' w), as the reference pairs its latest
// positions with its last cell lists.
sph_status twoscale_sub_step(Engine& e) {
    if (!twoscale_steady(e)) return e.fail(SPH_INVALID_ARG, "sub-particle step needs a two-scale step first");
    TwoScale& t = *e.ts;
    hipStream_t st = e.stream;
    e.mark(kMarkSubStep);
    hipLaunchKernelGGL(k_ts_unsort, dim3(nblocks(t.cap)), dim3(kB), 0, st, t.counts, t.sperm, t.spos_out, t.svel_out,
                       t.spos, t.svel);
    hipLaunchKernelGGL(k_ts_project, dim3(kProjGrid), dim3(kB), 0, st, t.counts, t.spos, t.svel, e.pos_a, e.pos_a,
                       e.range, e.d_scenes);
    e.mark(kStageSubStepProject);
    sub_pass(e, t, false);
    return e.check_launch("sub-particle step");
}

sph_status twoscale_export(Engine& e, float4* out, sph_render_view* view) {
    TwoScale& t = *e.ts;
    hipLaunchKernelGGL(k_ts_export, dim3(nblocks(t.cap)), dim3(kB), 0, e.stream, t.counts, t.spos_out, t.svel_out,
                       t.sperm, t.sid, out);
    sph_status st = e.check_launch("sub-particle render export");
    if (st != SPH_OK) return st;
    TSCounts c;
    if (hipStreamSynchronize(e.stream) != hipSuccess ||
        hipMemcpy(&c, t.counts, sizeof(c), hipMemcpyDeviceToHost) != hipSuccess)
        return e.fail(SPH_DEVICE_ERROR, "count read failed");
    view->sub_positions = reinterpret_cast<const float*>(out);
    view->default_ids = t.default_ids;
    view->sub_ids = t.sid;
    view->num_default = c.ndefault;
    view->num_sub_regular = 8 * c.nsplit;
    view->num_sub_boundary = 8 * c.nboundary;
    return SPH_OK;
}

sph_status twoscale_counts(Engine& e, sph_counts* out) {
    TSCounts c;
    if (hipStreamSynchronize(e.stream) != hipSuccess ||
        hipMemcpy(&c, e.ts->counts, sizeof(c), hipMemcpyDeviceToHost) != hipSuccess)
        return e.fail(SPH_DEVICE_ERROR, "count read failed");
    out->num_split = c.nsplit;
    out->num_boundary = c.nboundary;
    out->num_default = c.ndefault;
    out->num_sub = c.nsub;
    return SPH_OK;
}

size_t twoscale_bytes(const Engine& e) { return e.ts ? e.ts->bytes : 0; }

uint8_t* twoscale_final_state(Engine& e) { return e.ts->state_fin; }

bool twoscale_steady(const Engine& e) { return e.ts && e.ts->sub_cells_valid; }

sph_status twoscale_read(Engine& e, sph_field f, void* dst, size_t nbytes) {
    TwoScale& t = *e.ts;
    TSCounts c;
    if (hipMemcpy(&c, t.counts, sizeof(c), hipMemcpyDeviceToHost) != hipSuccess)
        return e.fail(SPH_DEVICE_ERROR, "count read failed");
    const uint32_t ns = c.nsub;
    auto need = [&](size_t want) -> bool { return nbytes == want; };
    auto cp = [&](void* h, const void* d, size_t b) -> sph_status {
        if (b && hipMemcpy(h, d, b, hipMemcpyDeviceToHost) != hipSuccess)
            return e.fail(SPH_DEVICE_ERROR, "device-to-host copy failed");
        return SPH_OK;
    };
    sph_status st;
    std::vector<uint32_t> perm(ns);
    auto load_perm = [&]() { return cp(perm.data(), t.sperm, ns * sizeof(uint32_t)); };
    switch (f) {
    case SPH_FIELD_SUB_POSITION:
    case SPH_FIELD_SUB_VELOCITY:
    case SPH_FIELD_SUB_ACCEL: {
        if (!need(3 * (size_t)ns * sizeof(float))) return e.fail(SPH_INVALID_ARG, "size mismatch");
        const float4* src = f == SPH_FIELD_SUB_POSITION ? t.spos_out : (f == SPH_FIELD_SUB_VELOCITY ? t.svel_out : t.sacc);
        std::vector<float4> tmp(ns);
        if ((st = cp(tmp.data(), src, ns * sizeof(float4))) != SPH_OK || (st = load_perm()) != SPH_OK) return st;
        float* o = (float*)dst;
        for (uint32_t i = 0; i < ns; ++i) {
            uint32_t k = perm[i];
            o[3 * (size_t)k] = tmp[i].x;
            o[3 * (size_t)k + 1] = tmp[i].y;
            o[3 * (size_t)k + 2] = tmp[i].z;
        }
        return SPH_OK;
    }
    case SPH_FIELD_SUB_DENSITY:
    case SPH_FIELD_SUB_PRESSURE: {
        if (!need((size_t)ns * sizeof(float))) return e.fail(SPH_INVALID_ARG, "size mismatch");
        std::vector<float4> tmp(ns);
        if ((st = cp(tmp.data(), t.svel_s, ns * sizeof(float4))) != SPH_OK || (st = load_perm()) != SPH_OK) return st;
        float* o = (float*)dst;
        const sph_params& P = e.params[0];
        for (uint32_t i = 0; i < ns; ++i) {
            float rho = tmp[i].w;
            o[perm[i]] = f == SPH_FIELD_SUB_DENSITY ? rho : P.gas_stiffness * (rho - P.rest_density);
        }
        return SPH_OK;
    }
    case SPH_FIELD_SUB_IDS:
        if (!need((size_t)ns * sizeof(int32_t))) return e.fail(SPH_INVALID_ARG, "size mismatch");
        return cp(dst, t.sid, ns * sizeof(uint32_t));
    case SPH_FIELD_SUB_SORTED_KEYS:
        if (!need((size_t)ns * sizeof(int32_t))) return e.fail(SPH_INVALID_ARG, "size mismatch");
        return cp(dst, t.sskey, ns * sizeof(uint32_t));
    case SPH_FIELD_SUB_SORTED_IDS: {
        if (!need((size_t)ns * sizeof(int32_t))) return e.fail(SPH_INVALID_ARG, "size mismatch");
        std::vector<uint32_t> sid(ns);
        if ((st = cp(sid.data(), t.sid, ns * sizeof(uint32_t))) != SPH_OK || (st = load_perm()) != SPH_OK) return st;
        uint32_t* o = (uint32_t*)dst;
        for (uint32_t i = 0; i < ns; ++i) o[i] = sid[perm[i]];
        return SPH_OK;
    }
    case SPH_FIELD_SUB_CELL_START:
    case SPH_FIELD_SUB_CELL_END: {
        if (!need((size_t)t.sub_cells * sizeof(int32_t))) return e.fail(SPH_INVALID_ARG, "size mismatch");
        std::vector<int2> tmp(t.sub_cells);
        if ((st = cp(tmp.data(), t.srange, (size_t)t.sub_cells * sizeof(int2))) != SPH_OK) return st;
        int32_t* o = (int32_t*)dst;
        for (uint32_t q = 0; q < t.sub_cells; ++q) {
            const int v = f == SPH_FIELD_SUB_CELL_START ? tmp[q].x : tmp[q].y;
            o[q] = v < 0 ? -1 : v;
        }
        return SPH_OK;
    }
    case SPH_FIELD_SUB_INPUT_POSITION:
    case SPH_FIELD_SUB_INPUT_VELOCITY: {
        if (!need(3 * (size_t)ns * sizeof(float))) return e.fail(SPH_INVALID_ARG, "size mismatch");
        std::vector<float4> tmp(ns);
        const float4* src = f == SPH_FIELD_SUB_INPUT_POSITION ? t.spos : t.svel;
        if ((st = cp(tmp.data(), src, ns * sizeof(float4))) != SPH_OK) return st;
        float* o = (float*)dst;
        for (uint32_t k = 0; k < ns; ++k) {
            o[3 * (size_t)k] = tmp[k].x;
            o[3 * (size_t)k + 1] = tmp[k].y;
            o[3 * (size_t)k + 2] = tmp[k].z;
        }
        return SPH_OK;
    }
    case SPH_FIELD_SPLIT_IDS:
        if (!need((size_t)c.nsplit * sizeof(int32_t))) return e.fail(SPH_INVALID_ARG, "size mismatch");
        return cp(dst, t.split_ids, c.nsplit * sizeof(uint32_t));
    case SPH_FIELD_BOUNDARY_IDS:
        if (!need((size_t)c.nboundary * sizeof(int32_t))) return e.fail(SPH_INVALID_ARG, "size mismatch");
        return cp(dst, t.boundary_ids, c.nboundary * sizeof(uint32_t));
    case SPH_FIELD_DEFAULT_IDS:
        if (!need((size_t)c.ndefault * sizeof(int32_t))) return e.fail(SPH_INVALID_ARG, "size mismatch");
        return cp(dst, t.default_ids, c.ndefault * sizeof(uint32_t));
    default:
        return e.fail(SPH_INVALID_ARG, "field not available");
    }
}

}  // namespace sph
