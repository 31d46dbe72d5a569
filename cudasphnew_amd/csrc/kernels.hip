// kernels.hip -- the coarse SPH step on gfx950.
//
// Data layout: particles live in SoA float4 streams that are re-ordered by
// cell every step, so that neighbour gathers are coalesced and every
// 3-cell x-row of a neighbourhood is one contiguous run of slots.
//   pos (x, y, z, density), vel (vx, vy, vz, density), acc (ax, ay, az, 0);
//   the density is the one of the step that wrote the record (0 before the
//   first step).
// The reference keeps AoS arrays in original-id order and gathers through
// the sorted id list instead (particle_simulation.cu:206,327).
//
// Per step:  [hash, first step only] -> clear last step's cells ->
//   sort by key (counting sort over the cell table, which also writes the
//   cell ranges; or radix sort + cell ranges) -> reorder (+ tie fix-up to
//   original-id order) -> density -> force + integrate + collision + next
//   key.
#include "kernels.h"

namespace sph {

namespace {

constexpr int kBlock = 256;

inline unsigned blocks_for(uint32_t n, int block) { return (n + block - 1) / block; }

// Scene record of a slot.  ONE (a single scene, every handle's case): index
// 0, a uniform address, so the record's fields are scalar loads; else the
// binary search over the scene begins, per lane (vector loads).
template <bool ONE>
__device__ __forceinline__ const DevScene& scene_for(const DevScene* __restrict__ scenes,
                                                     const uint32_t* __restrict__ begins, int nscenes, uint32_t s) {
    return ONE ? scenes[0] : scenes[scene_of(begins, nscenes, s)];
}

// Slot of a lane in the hit-mask kernels.  ONE: consecutive slots, lanes
// past the end repeat the last slot (not live).  Batch: the work list holds
// one entry per wavefront (x = first slot, y = scene << 7 | slot count
// 1..64), each inside one scene, so the wavefront's scene index is uniform
// and its record is read with scalar loads.  The entries are built once per
// batch from the scene ranges (the segmented sort keeps them fixed): first
// the 64-slot-aligned wavefronts of the whole slot range, each running the
// lanes of the scene its first slot belongs to, then, for the lanes of the
// later scenes that start inside one of them, one entry each -- at most one
// per scene, and every other wavefront's memory accesses stay aligned.
// Returns false for a wavefront with no work.
struct WaveSlot {
    uint32_t s;
    bool live;
    int scene;
};

template <int B, bool ONE>
__device__ __forceinline__ bool wave_slot(uint32_t n, const uint2* __restrict__ tiles, uint32_t ntiles,
                                          WaveSlot& w) {
    const uint32_t lane = threadIdx.x & 63u;
    if (ONE) {
        const uint32_t s0 = xcd_block(blockIdx.x, gridDim.x) * B + threadIdx.x;
        if ((s0 & ~63u) >= n) return false;
        w.live = s0 < n;
        w.s = w.live ? s0 : n - 1;
        w.scene = 0;
        return true;
    }
    const uint32_t t =
        __builtin_amdgcn_readfirstlane(xcd_block(blockIdx.x, gridDim.x) * (B / 64) + (threadIdx.x >> 6));
    if (t >= ntiles) return false;
    const uint2 e = tiles[t];
    const uint32_t cnt = e.y & 127u;
    w.scene = (int)(e.y >> 7);
    w.live = lane < cnt;
    w.s = e.x + (w.live ? lane : cnt - 1u);
    return true;
}

// compute_particle_hash (particle_simulation.cu:89-119) for the storage
// order.  Used before the first step; later keys come from the force kernel.
__global__ __launch_bounds__(kBlock) void k_hash(const float4* __restrict__ pos, uint32_t* __restrict__ key,
                                                 uint32_t n, const DevScene* __restrict__ scenes,
                                                 const uint32_t* __restrict__ begins, int nscenes,
                                                 uint32_t* __restrict__ status) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const DevScene& S = scenes[scene_of(begins, nscenes, i)];
    float4 p = pos[i];
    uint32_t flags = 0;
    key[i] = cell_key(S, p.x, p.y, p.z, flags);
    if (flags) atomicOr(status, flags);
}

// Invalidate the cells occupied last step (only those: the reference clears
// the whole table every step, cu:2362-2363).
__global__ __launch_bounds__(kBlock) void k_cells_clear(const uint32_t* __restrict__ skey, uint32_t n,
                                                        int2* __restrict__ range) {
    uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    uint32_t k = skey[s];
    if (s == 0 || skey[s - 1] != k) range[k] = make_int2(-1, -1);
}

// compute_cell_start_end (cu:155-193): start = first sorted slot of a key,
// end = one past the last.
__global__ __launch_bounds__(kBlock) void k_cells(const uint32_t* __restrict__ skey, uint32_t n,
                                                  int2* __restrict__ range) {
    uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    uint32_t k = skey[s];
    int* r = reinterpret_cast<int*>(range);
    if (s == 0 || skey[s - 1] != k) r[2 * (size_t)k] = (int)s;
    if (s == n - 1 || skey[s + 1] != k) r[2 * (size_t)k + 1] = (int)(s + 1);
}

// Gather storage -> sorted order.  The sort leaves each cell's particles in
// arrival (counting sort) or storage (radix sort) order; the reference sorts
// (hash, original id), so inside each cell run the slot is the rank of the
// original id (runs are short: a cell holds ~8 particles at rest density).
// SHUFFLE: a run that lies inside the lane's wavefront is ranked with
// shuffles of the lanes' ids, in a loop as long as the wavefront's longest
// such run, instead of reading the run's ids (a run crossing the
// wavefront's edge still reads them).  Measured: 82.8 -> 76.2 us for the
// 64-scene batch (cells of up to 17 particles), but 15.7 -> 17.6 us for the
// 1M scene (8 per cell, the reads hit the cache), so batches use it.
template <bool SHUFFLE>
__global__ __launch_bounds__(kBlock) void k_reorder(
    const uint32_t* __restrict__ skey, const uint32_t* __restrict__ perm, const int2* __restrict__ range,
    const uint32_t* __restrict__ oid_a, const float4* __restrict__ pos_a, const float4* __restrict__ vel_a,
    const uint8_t* __restrict__ state_a, uint32_t n, float4* __restrict__ pos_b, float4* __restrict__ vel_b,
    uint32_t* __restrict__ oid_b, uint8_t* __restrict__ state_b, uint32_t* __restrict__ slot_of_oid) {
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t k, src, me, rank = 0;
    int2 r;
    if (SHUFFLE) {
        const uint32_t w0 = s & ~63u;   // slot of the wavefront's first lane
        if (w0 >= n) return;            // the whole wavefront lies past the end
        const bool valid = s < n;
        k = valid ? skey[s] : 0u;
        r = valid ? range[k] : make_int2(0, 0);
        src = valid ? perm[s] : 0u;
        me = valid ? oid_a[src] : 0u;
        const bool inwave = valid && (uint32_t)r.x >= w0 && (uint32_t)r.y <= w0 + 64u;
        const int len = inwave ? r.y - r.x : 0;
        const int maxlen = __builtin_amdgcn_readfirstlane(__ockl_wfred_max_i32(len));
        const int first = inwave ? r.x - (int)w0 : 0;   // lane of the run's first slot
        for (int j = 0; j < maxlen; ++j) {   // every lane of the wavefront takes part
            const uint32_t o = __shfl(me, first + min(j, max(len - 1, 0)), 64);
            rank += (j < len && o < me) ? 1u : 0u;
        }
        if (!valid) return;
        if (!inwave)
            for (int q = r.x; q < r.y; ++q) rank += (oid_a[perm[q]] < me) ? 1u : 0u;
    } else {
        if (s >= n) return;
        k = skey[s];
        r = range[k];
        src = perm[s];
        me = oid_a[src];
        for (int q = r.x; q < r.y; ++q) rank += (oid_a[perm[q]] < me) ? 1u : 0u;
    }
    const uint32_t dst = (uint32_t)r.x + rank;
    pos_b[dst] = pos_a[src];
    vel_b[dst] = vel_a[src];
    oid_b[dst] = me;
    if (state_a) state_b[dst] = state_a[src];
    if (slot_of_oid) slot_of_oid[me] = dst;
}

// Positions by original id for a renderer: (x, y, z, density).
__global__ __launch_bounds__(kBlock) void k_export_positions(uint32_t n, const float4* __restrict__ pos,
                                                             const float4* __restrict__ vel,
                                                             const uint32_t* __restrict__ oid,
                                                             float4* __restrict__ out) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float4 p = pos[i];
    out[oid[i]] = make_float4(p.x, p.y, p.z, vel[i].w);
}

// Contiguous slot run of the cells [i0, i1] of one x-row (cells of a row
// have consecutive keys, so their particles are consecutive slots).
__device__ __forceinline__ void row_run(const int2* __restrict__ range, uint32_t rowbase, int i0, int i1, int& b,
                                        int& e) {
    b = 0x7fffffff;
    e = -1;
    for (int i = i0; i <= i1; ++i) {
        int2 r = range[rowbase + (uint32_t)i];
        if (r.x >= 0) {
            b = min(b, r.x);
            e = max(e, r.y);
        }
    }
}

// Neighbour range of a particle on the coarse grid (compute_grid_coordinate,
// cu:1238-1258) and the nine x-row runs of it.
struct Rows {
    int lx, hx, ly, hy, lz, hz;
};

__device__ __forceinline__ Rows neighbour_rows(const DevScene& S, float4 xi) {
    const float h = S.h;
    Rows r;
    r.lx = range_coord(xi.x, -h, S.ox, S.cell, S.nx); r.hx = range_coord(xi.x, h, S.ox, S.cell, S.nx);
    r.ly = range_coord(xi.y, -h, S.oy, S.cell, S.ny); r.hy = range_coord(xi.y, h, S.oy, S.cell, S.ny);
    r.lz = range_coord(xi.z, -h, S.oz, S.cell, S.nz); r.hz = range_coord(xi.z, h, S.oz, S.cell, S.nz);
    return r;
}

// Slot run of row (dk, dj) of the 3x3 row set; empty when the row is outside
// the clamped range (e.g. at the grid edge, or the reference's wrapped upper
// coordinate).
__device__ __forceinline__ void row_slots(const int2* __restrict__ range, const DevScene& S, const Rows& R, int dk,
                                          int dj, float4 xi, int& b, int& e) {
    const int k = R.lz + dk, j = R.ly + dj;
    b = 0;
    e = 0;
    if (k > R.hz || j > R.hy) return;
    // a row whose y-z cell rectangle is farther than h from the particle
    // holds no neighbour (its candidates would add exact zeros); the 1e-5
    // margin covers the rounding of the cell bounds and of the keys
    const float y0 = S.oy + (float)j * S.cell, z0 = S.oz + (float)k * S.cell;
    const float dy = fmaxf(fmaxf(y0 - xi.y, xi.y - (y0 + S.cell)), 0.0f);
    const float dz = fmaxf(fmaxf(z0 - xi.z, xi.z - (z0 + S.cell)), 0.0f);
    if ((dy * dy + dz * dz) * 0.99999f > S.h2) return;
    uint32_t rowbase = S.cell_offset + (uint32_t)(S.nx * (S.ny * k + j));
    row_run(range, rowbase, R.lx, R.hx, b, e);
    if (e < b) { b = 0; e = 0; }
}

// Pair terms of the acceleration (cu:1486-1555) accumulated without the
// constant factors: m, gradSpiky, mu, laplVisc, gradPoly6, laplPoly6 are
// applied once in force_finish.
struct PairSums {
    float fpx, fpy, fpz, fvx, fvy, fvz, cgx, cgy, cgz, cl;
};

// GUARD: the caller has not tested r2 > 0; a pair at distance 0 then adds
// exact zeros (inverse distance and 1/rho_j taken as 0).
template <bool GUARD = false>
__device__ __forceinline__ void pair_terms(PairSums& a, float4 xi, float4 vi, float rhoi, float Ai, float h, float h2,
                                           float stiff, float kr0, float4 xj, float4 vj) {
    float dx = xi.x - xj.x, dy = xi.y - xj.y, dz = xi.z - xj.z;
    float r2 = dx * dx + dy * dy + dz * dz;
    float inv_r = __builtin_amdgcn_rsqf(r2);
    float invj = __builtin_amdgcn_rcpf(vj.w);
    if (GUARD) {
        const bool ok = r2 > 0.0f;
        inv_r = ok ? inv_r : 0.0f;
        invj = ok ? invj : 0.0f;
    }
    float rn = r2 * inv_r;
    float pj = stiff * vj.w - kr0;
    float hr = h - rn;
    float qd = h2 - r2;
    float cp = (Ai + rhoi * pj * invj * invj) * hr * hr * inv_r;
    a.fpx += cp * dx; a.fpy += cp * dy; a.fpz += cp * dz;
    float wv = invj * hr;
    a.fvx += wv * (vj.x - vi.x); a.fvy += wv * (vj.y - vi.y); a.fvz += wv * (vj.z - vi.z);
    float wc = invj * qd * qd;
    a.cgx += wc * dx; a.cgy += wc * dy; a.cgz += wc * dz;
    a.cl += invj * qd * (3.0f * h2 - 7.0f * r2);
}

// Wall penalty + wall viscosity or surface tension (cu:407-466), a = F/rho - g
// (cu:470-473), symplectic Euler (integrate_euler, cu:632-665), the
// grid-driven collision response (collision_handling, cu:742-834) and the
// next step's cell key; writes the next storage order (this step's sorted
// order).
__device__ __forceinline__ void force_finish(const DevScene& S, uint32_t s, float4 xi, float4 vi, const PairSums& a,
                                             const float4* __restrict__ grid4, const float* __restrict__ gridd,
                                             const uint32_t* __restrict__ oid_b, const uint8_t* __restrict__ state_b,
                                             float4* __restrict__ acc, float4* __restrict__ pos_out,
                                             float4* __restrict__ vel_out, uint32_t* __restrict__ oid_out,
                                             uint8_t* __restrict__ state_out, uint32_t* __restrict__ key_out,
                                             uint32_t* __restrict__ status, const NeuralRefs& nr) {
    const float rhoi = vi.w;
    const float cpf = -S.mass * S.grad_spiky, cvf = S.mu * S.mass * S.lapl_visc;
    float Fx = cvf * a.fvx + cpf * a.fpx, Fy = cvf * a.fvy + cpf * a.fpy, Fz = cvf * a.fvz + cpf * a.fpz;
    const int wm = wall_mode(nr, s);
    float tx, ty, tz;
    texel_coords(S, xi.x, xi.y, xi.z, tx, ty, tz);
    // distance and wall viscosity from the grid (the wall-term switch
    // distWall < r0 and the penalty magnitude need its exact values); a cell
    // far from every surface takes no wall term and no lookup; with the
    // learned source the penalty and collision normals come from the
    // network's gradient instead of central differences of the grid
    float distWall = S.r0, wvisc = 0.0f;
    float4 r1 = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    if (wm != kWallNone) {
        float4 g = lerp3<float4>(grid4, S, tex_setup(S, tx, ty, tz));
        distWall = -g.x;
        wvisc = g.z;
        if (wm == kWallNeural && nr.need[s]) r1 = nr.rec[s];   // central-difference gradient of the learned distance at x
    }
    // the network's normal exists where the density pass asked for it
    const bool net_normal = wm == kWallNeural && nr.need[s];
    if (distWall < S.r0) {
        float coeff = rhoi / (S.dt * S.dt) * (S.r0 - distWall);
        // normalized-coordinate gradient: learned source dD/dx times
        // dx/du = L = dims / bscale per axis
        float3 gn = net_normal ? make_float3(r1.x * ((float)S.bnx / S.bscale_x), r1.y * ((float)S.bny / S.bscale_y),
                                             r1.z * ((float)S.bnz / S.bscale_z))
                               : dist_grad_normalized(gridd, S, tx, ty, tz);
        Fx -= coeff * gn.x + vi.x * wvisc;
        Fy -= coeff * gn.y + vi.y * wvisc;
        Fz -= coeff * gn.z + vi.z * wvisc;
    } else {
        float cm = S.mass * S.grad_poly6;
        float gx = cm * a.cgx, gy = cm * a.cgy, gz = cm * a.cgz;
        float gnorm = sqrtf(gx * gx + gy * gy + gz * gz);
        if (gnorm > S.norm_thresh) {
            float fc = S.sigma * (S.mass * S.lapl_poly6 * a.cl) / gnorm;
            Fx -= fc * gx;
            Fy -= fc * gy;
            Fz -= fc * gz;
        }
    }
    float ax = Fx / rhoi, ay = Fy / rhoi - S.gravity, az = Fz / rhoi;
    acc[s] = make_float4(ax, ay, az, 0.0f);
    float vx = vi.x + S.dt * ax, vy = vi.y + S.dt * ay, vz = vi.z + S.dt * az;
    float px = xi.x + S.dt * vx, py = xi.y + S.dt * vy, pz = xi.z + S.dt * vz;
    if (S.collision && wm != kWallNone) {
        float D;
        float3 gw;
        float cx, cy, cz;
        texel_coords(S, px, py, pz, cx, cy, cz);
        D = dist_at(gridd, S, cx, cy, cz);
        if (D > 0.0f) {
            if (net_normal) {
                gw = make_float3(r1.x, r1.y, r1.z);   // learned source: its gradient at the old position
            } else {
                gw = dist_grad_normalized(gridd, S, cx, cy, cz);
                // normalized-coordinate gradient -> world gradient: divide by L
                gw.x *= S.bscale_x * (1.0f / (float)S.bnx);
                gw.y *= S.bscale_y * (1.0f / (float)S.bny);
                gw.z *= S.bscale_z * (1.0f / (float)S.bnz);
            }
        }
        if (D > 0.0f) {
            float gl = sqrtf(gw.x * gw.x + gw.y * gw.y + gw.z * gw.z);
            if (gl > 0.0f) {
                float nx_ = gw.x / gl, ny_ = gw.y / gl, nz_ = gw.z / gl;
                float vn = sqrtf(vx * vx + vy * vy + vz * vz);
                float dp = nx_ * vx + ny_ * vy + nz_ * vz;
                float c = 1.0f;
                if (S.restitution != 0.0f && vn > 0.0f) c += S.restitution * D / (S.dt * vn);
                c *= dp;
                vx -= c * nx_; vy -= c * ny_; vz -= c * nz_;
                px -= D * nx_; py -= D * ny_; pz -= D * nz_;
            }
        }
    }
    uint32_t flags = 0;
    uint32_t kk = cell_key(S, px, py, pz, flags);
    if (!(isfinite(px) && isfinite(py) && isfinite(pz) && isfinite(vx) && isfinite(vy) && isfinite(vz)))
        flags |= kStatusNonFinite;
    if (flags) atomicOr(status, flags);
    pos_out[s] = make_float4(px, py, pz, rhoi);
    vel_out[s] = make_float4(vx, vy, vz, rhoi);
    key_out[s] = kk;
    oid_out[s] = oid_b[s];
    if (state_out) state_out[s] = state_b[s];
}

__device__ __forceinline__ float wall_density(const DevScene& S, const float4* __restrict__ grid4, float4 xi,
                                              const NeuralRefs& nr, uint32_t s) {
    // the density pass flags the network's particles; it also starts their
    // lists (the network pass appends after it)
    if (s == 0 && nr.list_count)
        for (int i = 0; i < nr.nscenes; ++i) nr.list_count[i] = 0u;
    const int wm = wall_mode(nr, s);
    if (wm == kWallNone) return 0.0f;
    float tx, ty, tz;
    texel_coords(S, xi.x, xi.y, xi.z, tx, ty, tz);
    const float4 g = lerp3<float4>(grid4, S, tex_setup(S, tx, ty, tz));
    // learned source: the force pass needs the network's normal exactly
    // where the wall penalty acts (distWall = -D < r0)
    if (wm == kWallNeural && nr.need) nr.need[s] = g.x > -S.r0;
    return g.y;
}

// compute_particle_density_pressure (cu:196-256, 1385-1419), one thread per
// particle.  Density goes to vel.w; pressure k(rho - rho0) is recomputed
// where needed.
template <int B, bool ONE>
__global__ __launch_bounds__(B) void k_density(const float4* __restrict__ pos, float4* __restrict__ vel,
                                                    const int2* __restrict__ range, uint32_t n,
                                                    const DevScene* __restrict__ scenes,
                                                    const uint32_t* __restrict__ begins, int nscenes,
                                                    const float4* __restrict__ grid4, NeuralRefs nr) {
    uint32_t s = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    if (s >= n) return;
    const DevScene& S = scene_for<ONE>(scenes, begins, nscenes, s);
    float4 xi = pos[s];
    const float h2 = S.h2;
    Rows R = neighbour_rows(S, xi);
    float sum = 0.0f;
    for (int dk = 0; dk < 3; ++dk)
        for (int dj = 0; dj < 3; ++dj) {
            int b, e;
            row_slots(range, S, R, dk, dj, xi, b, e);
#pragma unroll 4
            for (int q = b; q < e; ++q) {
                float4 xj = pos[q];
                float dx = xj.x - xi.x, dy = xj.y - xi.y, dz = xj.z - xi.z;
                float r2 = dx * dx + dy * dy + dz * dz;
                float d = fmaxf(h2 - r2, 0.0f);
                sum += d * d * d;
            }
        }
    float rho = sum * (S.poly6 * S.mass) + wall_density(S, grid4, xi, nr, s);
    reinterpret_cast<float*>(vel)[4 * (size_t)s + 3] = rho;
}

// Wavefront-cooperative density: a block owns 256 consecutive sorted
// particles and walks them in L rounds; in a round, L adjacent lanes share one
// particle and take every L-th candidate of its row runs.  Lanes of a
// particle read consecutive candidates and the particles of a wavefront sit
// in the same or adjacent cells, so one load instruction touches a few cache
// lines instead of one run per lane.  The L partial sums meet through
// cross-lane shuffles; the per-particle epilogue runs one thread per particle.
template <int L>
__global__ __launch_bounds__(kBlock) void k_density_coop(const float4* __restrict__ pos, float4* __restrict__ vel,
                                                         const int2* __restrict__ range, uint32_t n,
                                                         const DevScene* __restrict__ scenes,
                                                         const uint32_t* __restrict__ begins, int nscenes,
                                                         const float4* __restrict__ grid4, NeuralRefs nr) {
    constexpr int PR = kBlock / L;
    __shared__ float s_sum[kBlock];
    const uint32_t P0 = xcd_block(blockIdx.x, gridDim.x) * kBlock;
    const int c = threadIdx.x % L;
    for (int round = 0; round < L; ++round) {
        const int pl = round * PR + threadIdx.x / L;
        const uint32_t s = P0 + pl;
        float sum = 0.0f;
        if (s < n) {
            const DevScene& S = scenes[scene_of(begins, nscenes, s)];
            const float4 xi = pos[s];
            const float h2 = S.h2;
            Rows R = neighbour_rows(S, xi);
            for (int dk = 0; dk < 3; ++dk)
                for (int dj = 0; dj < 3; ++dj) {
                    int b, e;
                    row_slots(range, S, R, dk, dj, xi, b, e);
                    for (int q = b + c; q < e; q += L) {
                        float4 xj = pos[q];
                        float dx = xj.x - xi.x, dy = xj.y - xi.y, dz = xj.z - xi.z;
                        float r2 = dx * dx + dy * dy + dz * dz;
                        float d = fmaxf(h2 - r2, 0.0f);
                        sum += d * d * d;
                    }
                }
        }
#pragma unroll
        for (int o = 1; o < L; o <<= 1) sum += __shfl_xor(sum, o, 64);
        if (c == 0) s_sum[pl] = sum;
    }
    __syncthreads();
    const uint32_t s = P0 + threadIdx.x;
    if (s >= n) return;
    const DevScene& S = scenes[scene_of(begins, nscenes, s)];
    const float4 xi = pos[s];
    float rho = s_sum[threadIdx.x] * (S.poly6 * S.mass) + wall_density(S, grid4, xi, nr, s);
    reinterpret_cast<float*>(vel)[4 * (size_t)s + 3] = rho;
}

// compute_particle_acceleration_ifsurf (cu:315-474, 1459-1556) + integration,
// one thread per particle.  Per run of up to 64 candidates a cheap distance
// test records the hits in a 64-bit mask, then the pair terms run for the
// set bits only: roughly one candidate in six is within h, so a fused loop
// would run the pair body for almost every candidate of a wavefront.
template <int B, bool ONE>
__global__ __launch_bounds__(B) void k_force(
    const float4* __restrict__ pos, const float4* __restrict__ vel, const int2* __restrict__ range, uint32_t n,
    const DevScene* __restrict__ scenes, const uint32_t* __restrict__ begins, int nscenes,
    const float4* __restrict__ grid4, const float* __restrict__ gridd, const uint32_t* __restrict__ oid_b,
    const uint8_t* __restrict__ state_b, float4* __restrict__ acc, float4* __restrict__ pos_out,
    float4* __restrict__ vel_out, uint32_t* __restrict__ oid_out, uint8_t* __restrict__ state_out,
    uint32_t* __restrict__ key_out, uint32_t* __restrict__ status, NeuralRefs nr) {
    uint32_t s = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    if (s >= n) return;
    const DevScene& S = scene_for<ONE>(scenes, begins, nscenes, s);
    const float4 xi = pos[s];
    const float4 vi = vel[s];
    const float rhoi = vi.w;
    const float h = S.h, h2 = S.h2;
    const float Ai = S.stiff * (rhoi - S.rho0) / rhoi;   // p_i / rho_i
    const float kr0 = S.stiff * S.rho0;
    Rows R = neighbour_rows(S, xi);
    PairSums a = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (int dk = 0; dk < 3; ++dk)
        for (int dj = 0; dj < 3; ++dj) {
            int b, e;
            row_slots(range, S, R, dk, dj, xi, b, e);
            for (int base = b; base < e; base += 64) {
                const int cnt = min(64, e - base);
                uint64_t hits = 0;
#pragma unroll 4
                for (int t = 0; t < cnt; ++t) {
                    float4 xj = pos[base + t];
                    float dx = xi.x - xj.x, dy = xi.y - xj.y, dz = xi.z - xj.z;
                    float r2 = dx * dx + dy * dy + dz * dz;
                    hits |= (uint64_t)(r2 <= h2 && r2 > 0.0f) << t;
                }
                while (hits) {
                    const int t = __ffsll((unsigned long long)hits) - 1;
                    hits &= hits - 1;
                    const int q = base + t;
                    pair_terms(a, xi, vi, rhoi, Ai, h, h2, S.stiff, kr0, pos[q], vel[q]);
                }
            }
        }
    force_finish(S, s, xi, vi, a, grid4, gridd, oid_b, state_b, acc, pos_out, vel_out, oid_out, state_out, key_out,
                 status, nr);
}

// Wavefront-cooperative force.  Per round, L adjacent lanes share a particle:
// they split its candidates, ballot the hits and append them to the
// particle's hit list in LDS (one wavefront-wide vote per candidate batch,
// no atomics), then split the list for the pair terms, so the expensive body
// runs on dense hits only.  The ten partial sums meet through shuffles and
// the epilogue runs one thread per particle.  A particle with more than CAP
// hits (strong compression) is finished by its first lane with the direct
// loop.
template <int L>
__global__ __launch_bounds__(kBlock) void k_force_coop(
    const float4* __restrict__ pos, const float4* __restrict__ vel, const int2* __restrict__ range, uint32_t n,
    const DevScene* __restrict__ scenes, const uint32_t* __restrict__ begins, int nscenes,
    const float4* __restrict__ grid4, const float* __restrict__ gridd, const uint32_t* __restrict__ oid_b,
    const uint8_t* __restrict__ state_b, float4* __restrict__ acc, float4* __restrict__ pos_out,
    float4* __restrict__ vel_out, uint32_t* __restrict__ oid_out, uint8_t* __restrict__ state_out,
    uint32_t* __restrict__ key_out, uint32_t* __restrict__ status, NeuralRefs nr) {
    static_assert(L >= 2 && L <= 16 && (L & (L - 1)) == 0, "L must be 2, 4, 8 or 16");
    constexpr int PR = kBlock / L;
    constexpr int CAP = L == 2 ? 64 : 96;
    __shared__ int s_list[PR][CAP];
    __shared__ float s_acc[10][kBlock];
    const uint32_t P0 = xcd_block(blockIdx.x, gridDim.x) * kBlock;
    const int lane = threadIdx.x & 63;
    const int c = threadIdx.x % L;
    const int gshift = (lane / L) * L;              // first lane of the group
    const uint32_t below = (1u << c) - 1u;
    for (int round = 0; round < L; ++round) {
        const int gp = threadIdx.x / L;
        const int pl = round * PR + gp;
        const uint32_t s = P0 + pl;
        const bool valid = s < n;
        const DevScene& S = scenes[scene_of(begins, nscenes, valid ? s : 0)];
        const float4 xi = valid ? pos[s] : make_float4(0, 0, 0, 0);
        const float4 vi = valid ? vel[s] : make_float4(0, 0, 0, 1);
        const float h = S.h, h2 = S.h2;
        const float rhoi = vi.w;
        const float Ai = S.stiff * (rhoi - S.rho0) / rhoi;
        const float kr0 = S.stiff * S.rho0;
        Rows R = neighbour_rows(S, xi);
        int cnt = 0;
        // phase 1: distance tests, hits compacted into the particle's list
        for (int dk = 0; dk < 3; ++dk)
            for (int dj = 0; dj < 3; ++dj) {
                int b = 0, e = 0;
                if (valid) row_slots(range, S, R, dk, dj, xi, b, e);
                for (int q = b + c; __any(q < e); q += L) {
                    bool hit = false;
                    if (q < e) {
                        float4 xj = pos[q];
                        float dx = xi.x - xj.x, dy = xi.y - xj.y, dz = xi.z - xj.z;
                        float r2 = dx * dx + dy * dy + dz * dz;
                        hit = r2 <= h2 && r2 > 0.0f;
                    }
                    const uint64_t vote = __ballot(hit);
                    const uint32_t gbits = (uint32_t)(vote >> gshift) & ((1u << L) - 1u);
                    if (hit) {
                        int idx = cnt + __popc(gbits & below);
                        if (idx < CAP) s_list[gp][idx] = q;
                    }
                    cnt += __popc(gbits);
                }
            }
        __syncthreads();
        // phase 2: pair terms over the hit list, split across the lanes
        PairSums a = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
        const int m = min(cnt, CAP);
        for (int i = c; i < m; i += L) {
            const int q = s_list[gp][i];
            pair_terms(a, xi, vi, rhoi, Ai, h, h2, S.stiff, kr0, pos[q], vel[q]);
        }
        if (cnt > CAP && c == 0) {
            // overflow: redo this particle directly
            a = PairSums{0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
            for (int dk = 0; dk < 3; ++dk)
                for (int dj = 0; dj < 3; ++dj) {
                    int b, e;
                    row_slots(range, S, R, dk, dj, xi, b, e);
                    for (int q = b; q < e; ++q) {
                        float4 xj = pos[q];
                        float dx = xi.x - xj.x, dy = xi.y - xj.y, dz = xi.z - xj.z;
                        float r2 = dx * dx + dy * dy + dz * dz;
                        if (r2 <= h2 && r2 > 0.0f) pair_terms(a, xi, vi, rhoi, Ai, h, h2, S.stiff, kr0, xj, vel[q]);
                    }
                }
        }
        // shuffles run on every lane (uniform control flow); a particle that
        // overflowed keeps its first lane's direct sums
        float* av = &a.fpx;
#pragma unroll
        for (int t = 0; t < 10; ++t) {
            float v = av[t];
#pragma unroll
            for (int o = 1; o < L; o <<= 1) v += __shfl_xor(v, o, 64);
            av[t] = cnt <= CAP ? v : av[t];
        }
        if (c == 0) {
#pragma unroll
            for (int t = 0; t < 10; ++t) s_acc[t][pl] = av[t];
        }
        __syncthreads();
    }
    const uint32_t s = P0 + threadIdx.x;
    if (s >= n) return;
    const DevScene& S = scenes[scene_of(begins, nscenes, s)];
    PairSums a;
    float* av = &a.fpx;
#pragma unroll
    for (int t = 0; t < 10; ++t) av[t] = s_acc[t][threadIdx.x];
    force_finish(S, s, pos[s], vel[s], a, grid4, gridd, oid_b, state_b, acc, pos_out, vel_out, oid_out, state_out,
                 key_out, status, nr);
}

// Hit-mask pipeline.  The density pass already evaluates the distance of
// every candidate of the nine row runs; it records, per particle and row, the
// row's first slot and which of its candidates are force neighbours, packed
// in one 64-bit word per row (masks[row * n + slot], coalesced; the format is
// in common.h): near rest density (at most 32 candidates per row, 24 on a
// lattice) the word is all there is; the longer rows of a compressed flow
// add 32-bit masks for their further chunks of 32 candidates, and rows of
// more than 128 candidates are marked kRowLong.  The force pass then needs
// neither the distance tests nor the row runs (no cell-table reads, no
// neighbour-range divisions): it expands the masks into a per-thread hit
// list in LDS and runs the pair terms over the list, so a wavefront's trip
// count is the largest neighbour count of its lanes rather than the sum over
// rows of the largest per-row count.  kRowLong rows are scanned by the force
// pass directly.
//
// Mask bits: r^2 <= h^2, without the particle's own slot.  A distinct
// particle at distance 0 keeps its bit; the force pass gives it no
// contribution (the pair terms require 0 < r^2, as the reference's
// `rn > 0`).

// Neighbour bits of c <= 32 candidates from the sign bits of h2 - r2 shifted
// in newest-first (bit 0 = candidate c - 1, set = farther than h): reverse,
// invert, drop the unused high part.
__device__ __forceinline__ uint32_t near_bits(uint32_t far, int c) {
    return c > 0 ? __builtin_bitreverse32(~far) >> (32 - c) : 0u;
}

// One candidate of the density pass: poly6 sum and the "farther than h" bit.
// d = h2 - dz^2 - dy^2 - dx^2 as a chain of three fused multiply-adds (one
// instruction fewer than forming r^2 and subtracting it).  Each step rounds
// an exact result, which keeps its sign (and an exact zero is +0), so d's
// sign bit is set exactly when this d is negative, i.e. when the candidate's
// poly6 term is zero: the bit is that "farther than h" bit and one funnel
// shift per candidate builds the mask.
__device__ __forceinline__ void density_candidate(float xjx, float xjy, float xjz, float4 xi, float h2, float& sum,
                                                  uint32_t& far) {
    float dx = xjx - xi.x, dy = xjy - xi.y, dz = xjz - xi.z;
    float d = fmaf(-dx, dx, fmaf(-dy, dy, fmaf(-dz, dz, h2)));
    far = __builtin_amdgcn_alignbit(far, __float_as_uint(d), 31);
    d = fmaxf(d, 0.0f);
    sum += d * d * d;
}

// The candidates [0, cnt) of one row run: adds their poly6 terms to sum and
// returns the force-neighbour mask of candidates 0..31 (bit t = candidate t
// within h; self = the particle's own position in the run, or any value >=
// cnt).  Each further chunk k of 32 candidates is handed to put(k, mask).
// at(t) -> (x, y, z) of candidate t.
template <class At, class Put>
__device__ __forceinline__ uint32_t density_row(const At& at, int cnt, float4 xi, float h2, uint32_t self, float& sum,
                                                const Put& put) {
    uint32_t far = 0;
    const int c32 = min(cnt, 32);
#pragma unroll 4
    for (int t = 0; t < c32; ++t) {
        const float3 p = at(t);
        density_candidate(p.x, p.y, p.z, xi, h2, sum, far);
    }
    uint32_t m = near_bits(far, c32);
    if (self < (uint32_t)c32) m &= ~(1u << self);
    // rows of a compressed flow
    for (int c0 = 32; c0 < cnt; c0 += 32) {
        const int cc = min(cnt - c0, 32);
        uint32_t f = 0;
        for (int t = 0; t < cc; ++t) {
            const float3 p = at(c0 + t);
            density_candidate(p.x, p.y, p.z, xi, h2, sum, f);
        }
        uint32_t mc = near_bits(f, cc);
        const uint32_t sc = self - (uint32_t)c0;
        if (sc < (uint32_t)cc) mc &= ~(1u << sc);
        put((uint32_t)c0 >> 5, mc);
    }
    return m;
}

// Window of candidate positions per wavefront, in LDS (x, y, z planes).  The
// 64 particles of a wavefront are consecutive slots (a few neighbouring cells
// of one cell row, as a rule), so for each of the nine row offsets their
// runs lie in one short slot window [lo, hi).  The wavefront copies each
// window once with coalesced loads, and the candidate loop reads LDS: one
// global load per particle and row in place of one per particle and
// candidate, which leaves the pass bound by its arithmetic instead of the
// texture data path.  A row whose window does not fit in what is left of
// the wavefront's kWin entries (the first rows fill it; a wavefront that
// straddles two cell rows or two scenes needs more) reads global memory.
// kWin trades rows served from LDS against occupancy (the window is the
// kernel's LDS): measured at 1M particles, 768 entries (all nine rows, four
// wavefronts per SIMD) 131 us, 640 124 us, 512 121 us, 384 120 us, 256 135 us.
// Three planes rather than float4 entries: one 12-byte read per candidate
// without address arithmetic (float4 window) measured 114 us against 109 us
// at 384 entries (117 us at 320, 122 us at 256) -- the larger window costs
// more occupancy than the saved instructions gain.
constexpr int kWin = 384;

template <int B, bool ONE>
__global__ __launch_bounds__(B) void k_density_masks(const float4* __restrict__ pos, float4* __restrict__ vel,
                                                     const int2* __restrict__ range, uint32_t n,
                                                     const DevScene* __restrict__ scenes,
                                                     const uint2* __restrict__ tiles, uint32_t ntiles,
                                                     const float4* __restrict__ grid4,
                                                     uint64_t* __restrict__ masks, NeuralRefs nr,
                                                     SplitRefs split) {
    __shared__ float s_win[B / 64][3 * kWin];
    WaveSlot ws;
    if (!wave_slot<B, ONE>(n, tiles, ntiles, ws)) return;
    // every lane of a live wavefront takes part in the window copies; lanes
    // past the end repeat the last particle and store nothing
    const bool live = ws.live;
    const uint32_t s = ws.s;
    const int lane = threadIdx.x & 63;
    float* win = s_win[threadIdx.x >> 6];
    {
        const DevScene& S = scenes[ws.scene];
        const float4 xi = pos[s];
        const float h2 = S.h2;
        Rows R = neighbour_rows(S, xi);
        int rb[9], re[9], wofs[9];
        bool inlds[9];
        int fill = 0;
    #pragma unroll
        for (int row = 0; row < 9; ++row) {
            row_slots(range, S, R, row / 3, row % 3, xi, rb[row], re[row]);
            const bool any = re[row] > rb[row];
            const int lo = __builtin_amdgcn_readfirstlane(__ockl_wfred_min_i32(any ? rb[row] : 0x7fffffff));
            const int hi = __builtin_amdgcn_readfirstlane(__ockl_wfred_max_i32(any ? re[row] : -1));
            inlds[row] = hi > lo && hi - lo <= kWin - fill;
            wofs[row] = fill - lo;   // LDS index of slot q: q + wofs
            if (inlds[row]) {
                for (int i = lane; i < hi - lo; i += 64) {
                    const float4 p = pos[lo + i];
                    win[fill + i] = p.x;
                    win[kWin + fill + i] = p.y;
                    win[2 * kWin + fill + i] = p.z;
                }
                fill += hi - lo;
            }
        }
        // the window is private to the wavefront: its own LDS writes, in
        // program order, are visible to its later reads
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        float sum = 0.0f;
        uint32_t* const cm = chunk_masks(masks, n);
    #pragma unroll
        for (int row = 0; row < 9; ++row) {
            const int b = rb[row], cnt = re[row] - rb[row];
            // the particle's own slot is not a force neighbour
            const uint32_t self = s - (uint32_t)b;
            const bool keep = live && cnt <= kRowMaxCands;
            auto put = [&](uint32_t k, uint32_t mk) {
                if (keep) cm[chunk_at(n, k, row, s)] = mk;
            };
            uint32_t m;
            if (inlds[row]) {
                const float* w = win + (b + wofs[row]);
                m = density_row([&](int t) { return make_float3(w[t], w[kWin + t], w[2 * kWin + t]); }, cnt, xi, h2,
                                self, sum, put);
            } else {
                const F4Buf pb(pos);
                m = density_row([&](int t) { const float4 p = pb[b + t]; return make_float3(p.x, p.y, p.z); }, cnt,
                                xi, h2, self, sum, put);
            }
            const uint64_t w = (uint64_t)(uint32_t)b << 32 | m;
            if (live)
                masks[(size_t)row * n + s] =
                    cnt <= 32 ? w : (keep ? w | (uint64_t)((cnt - 1) >> 5) << 60 : kRowLong);
        }
        if (live) {
            float rho = sum * (S.poly6 * S.mass) + wall_density(S, grid4, xi, nr, s);
            reinterpret_cast<float*>(vel)[4 * (size_t)s + 3] = rho;
        }
        if (split.state) {
            // two-scale step: the particle's split flag (inside the split
            // box) on top of its shifted state, and the split particles'
            // slots appended to the split list (every lane takes part)
            bool in = false;
            if (live) {
                uint8_t st = (uint8_t)(split.state[s] << 2);
                in = xi.x >= S.split_lo[0] && xi.x <= S.split_hi[0] && xi.y >= S.split_lo[1] &&
                     xi.y <= S.split_hi[1] && xi.z >= S.split_lo[2] && xi.z <= S.split_hi[2];
                if (in) st |= 1;
                split.state[s] = st;
                split.state_fin[s] = st;
            }
            const uint64_t ins = __ballot(in);
            if (ins) {
                const int leader = __ffsll((unsigned long long)ins) - 1;
                uint32_t base = 0;
                if (lane == leader) base = atomicAdd(split.nmark, (uint32_t)__popcll(ins));
                base = __shfl(base, leader, 64);
                if (in) {
                    const uint64_t below = lane == 0 ? 0ull : (ins & (~0ull >> (64 - lane)));
                    split.mark_list[base + (uint32_t)__popcll(below)] = s;
                }
            }
        }
    }
}

// The hit list holds 16-bit entries (row << 7 | position in the row's run)
// and the nine run starts per thread: CAP * 512 + 9216 bytes of LDS per
// block, so that five blocks fit the LDS of a CU at CAP = 40 (four at 48,
// three at 64), and the occupancy hint asks the register allocation for as
// many wavefronts per SIMD.  The list is filled in rounds: the unexpanded
// hits stay in registers (the current chunk's mask of each row, the chunk
// index), a round expands up to CAP of them and runs the pair loop over
// them, and a lane whose round filled the list goes again -- so a compressed
// flow's longer lists keep the expansion loop free of pair arithmetic and
// the gathers pipelined.  Measured at 1M particles, hit list 48, on the
// dam's positions at steps 1000 and 2000 (rows of up to 53 and 88
// candidates, 12% and 82% of the particles with more than 48 neighbours):
// 286 -> 121 us and 760 -> 239 us against a single list with the overflow
// evaluated one gather at a time and rows past 32 candidates scanned; 87 ->
// 90 us near rest density.  The pair loop keeps the gathers of the next
// three neighbours in flight while it evaluates the current one.  Measured
// at 1M particles before the packed row words and the scalar scene loads:
// CAP 40 (5 wavefronts, 85 VGPRs) 127 us, 48 (4) 132 us, 32 at 6 wavefronts
// spills and takes 139 us (156 us with two gathers in flight instead of
// three).
template <int CAP, int B, bool ONE>
__global__ __launch_bounds__(B) __attribute__((amdgpu_waves_per_eu(CAP <= 40 ? 5 : (CAP <= 48 ? 4 : 3), 8))) void k_force_masks(
    const float4* __restrict__ pos, const float4* __restrict__ vel, const int2* __restrict__ range, uint32_t n,
    const DevScene* __restrict__ scenes, const uint2* __restrict__ tiles, uint32_t ntiles,
    const float4* __restrict__ grid4, const float* __restrict__ gridd, const uint32_t* __restrict__ oid_b,
    const uint8_t* __restrict__ state_b, float4* __restrict__ acc, float4* __restrict__ pos_out,
    float4* __restrict__ vel_out, uint32_t* __restrict__ oid_out, uint8_t* __restrict__ state_out,
    uint32_t* __restrict__ key_out, uint32_t* __restrict__ status, const uint64_t* __restrict__ masks,
    NeuralRefs nr) {
    // one column per thread: no barrier needed, consecutive threads hit
    // consecutive banks
    __shared__ uint16_t s_list[CAP][B];
    __shared__ int s_base[9][B];
    WaveSlot ws;
    if (!wave_slot<B, ONE>(n, tiles, ntiles, ws) || !ws.live) return;
    const uint32_t s = ws.s;
    {
        const DevScene& S = scenes[ws.scene];
        const int tid = threadIdx.x;
        // unexpanded hits: bits[row] = the rest of the current chunk's mask;
        // chunk / last: current and last chunk of each row, 2 bits per row
        uint32_t bits[9];
        uint32_t chunk = 0, last = 0, long_rows = 0;
        {
            uint64_t mk[9];
    #pragma unroll
            for (int row = 0; row < 9; ++row) mk[row] = masks[(size_t)row * n + s];
    #pragma unroll
            for (int row = 0; row < 9; ++row) {
                const bool lr = mk[row] == kRowLong;
                long_rows |= (uint32_t)lr << row;
                bits[row] = lr ? 0u : (uint32_t)mk[row];
                last |= lr ? 0u : row_last_chunk(mk[row]) << (2 * row);
                s_base[row][tid] = (int)row_first(mk[row]);
            }
        }
        const uint32_t* const cm = chunk_masks(masks, n);
        const float4 xi = pos[s];
        const float4 vi = vel[s];
        const float rhoi = vi.w;
        const float h = S.h, h2 = S.h2;
        const float Ai = S.stiff * (rhoi - S.rho0) / rhoi;   // p_i / rho_i
        const float kr0 = S.stiff * S.rho0;
        PairSums a = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
        const F4Buf pb(pos), vb(vel);
        int m;
        do {
            m = 0;
    #pragma unroll
            for (int row = 0; row < 9; ++row) {
                for (;;) {
                    const uint32_t c = chunk >> (2 * row) & 3u;
                    const int off = row << 7 | (int)(c << 5);
                    uint32_t lo = bits[row];
                    // a counted loop, unrolled
                    const int take = min((int)__popc(lo), CAP - m);
    #pragma unroll 8
                    for (int i = 0; i < take; ++i) {
                        const int t = __builtin_ctz(lo);
                        lo &= lo - 1;
                        s_list[m + i][tid] = (uint16_t)(off | t);
                    }
                    m += take;
                    bits[row] = lo;
                    if (lo || c == (last >> (2 * row) & 3u)) break;
                    // next chunk of a long row
                    chunk += 1u << (2 * row);
                    bits[row] = cm[chunk_at(n, c + 1, row, s)];
                }
            }
            if (m > 0) {
                // three entries in flight: each register set is reloaded with
                // the entry three places on right after its pair terms, so the
                // loop (unrolled by three by hand) moves no registers and waits
                // only for the oldest gathers.  Entries past the end repeat the
                // last one and are loaded but not evaluated.
                auto slot = [&](int k) {
                    const int ent = s_list[min(k, m - 1)][tid];
                    return s_base[ent >> 7][tid] + (ent & 127);
                };
                int q = slot(0);
                float4 x0 = pb[q], v0 = vb[q];
                q = slot(1);
                float4 x1 = pb[q], v1 = vb[q];
                q = slot(2);
                float4 x2 = pb[q], v2 = vb[q];
                int k = 0;
                for (; k + 3 < m; k += 3) {
                    pair_terms<true>(a, xi, vi, rhoi, Ai, h, h2, S.stiff, kr0, x0, v0);
                    q = slot(k + 3);
                    x0 = pb[q]; v0 = vb[q];
                    pair_terms<true>(a, xi, vi, rhoi, Ai, h, h2, S.stiff, kr0, x1, v1);
                    q = slot(k + 4);
                    x1 = pb[q]; v1 = vb[q];
                    pair_terms<true>(a, xi, vi, rhoi, Ai, h, h2, S.stiff, kr0, x2, v2);
                    q = slot(k + 5);
                    x2 = pb[q]; v2 = vb[q];
                }
                pair_terms<true>(a, xi, vi, rhoi, Ai, h, h2, S.stiff, kr0, x0, v0);
                if (k + 1 < m) pair_terms<true>(a, xi, vi, rhoi, Ai, h, h2, S.stiff, kr0, x1, v1);
                if (k + 2 < m) pair_terms<true>(a, xi, vi, rhoi, Ai, h, h2, S.stiff, kr0, x2, v2);
            }
        } while (m == CAP);   // a full list: hits may remain
        if (long_rows) {
            // rows of more than kRowMaxCands candidates: scanned here
            const Rows R = neighbour_rows(S, xi);
    #pragma unroll 1
            for (int row = 0; row < 9; ++row) {
                if (!(long_rows >> row & 1u)) continue;
                int b, e;
                row_slots(range, S, R, row / 3, row % 3, xi, b, e);
                for (int q = b; q < e; ++q) {
                    float4 xj = pos[q];
                    float dx = xi.x - xj.x, dy = xi.y - xj.y, dz = xi.z - xj.z;
                    float r2 = dx * dx + dy * dy + dz * dz;
                    if (r2 <= h2 && r2 > 0.0f) pair_terms<false>(a, xi, vi, rhoi, Ai, h, h2, S.stiff, kr0, xj, vel[q]);
                }
            }
        }
        force_finish(S, s, xi, vi, a, grid4, gridd, oid_b, state_b, acc, pos_out, vel_out, oid_out, state_out, key_out,
                     status, nr);
    }
}

// Surface particles (the reference's declared extractSurfaceParticles, on
// the positions of the last step's force pass): neighbour count including
// the particle itself and the offset of the neighbourhood's centre of mass,
// accumulated as differences x_j - x_i (no cancellation against |x|).
__global__ __launch_bounds__(kBlock) void k_surface(const float4* __restrict__ pos, const int2* __restrict__ range,
                                                    uint32_t n, const DevScene* __restrict__ scenes,
                                                    const uint32_t* __restrict__ begins, int nscenes,
                                                    const uint32_t* __restrict__ oid, float cm_thresh,
                                                    float n_thresh, uint8_t* __restrict__ out) {
    uint32_t s = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    if (s >= n) return;
    const DevScene& S = scenes[scene_of(begins, nscenes, s)];
    const float4 xi = pos[s];
    Rows R = neighbour_rows(S, xi);
    float sx = 0.0f, sy = 0.0f, sz = 0.0f, cnt = 1.0f;
    for (int dk = 0; dk < 3; ++dk)
        for (int dj = 0; dj < 3; ++dj) {
            int b, e;
            row_slots(range, S, R, dk, dj, xi, b, e);
            for (int q = b; q < e; ++q) {
                const float4 xj = pos[q];
                const float dx = xj.x - xi.x, dy = xj.y - xi.y, dz = xj.z - xi.z;
                const float r2 = dx * dx + dy * dy + dz * dz;
                if (r2 <= S.h2 && r2 > 0.0f) {
                    sx += dx;
                    sy += dy;
                    sz += dz;
                    cnt += 1.0f;
                }
            }
        }
    sx /= cnt;
    sy /= cnt;
    sz /= cnt;
    const float d = sqrtf(sx * sx + sy * sy + sz * sz);
    out[oid[s]] = (cnt < n_thresh || d > cm_thresh * S.h) ? 1 : 0;
}

// Boundary source of the cells of one scene from its grid: a cell whose
// particles can only reach texels holding the far-field values (distance
// -r0, no wall density or viscosity) takes no wall term and no lookup --
// exactly what the lookup would return.  The footprint is the cell's texel
// range widened by two texels on each side (the central differences of the
// penalty, the collision lookup after a move of less than a spacing).
__global__ void k_wall_modes(DevScene S, const float4* __restrict__ grid4, uint8_t near_value,
                             uint8_t* __restrict__ mode) {
    const uint32_t ncell = (uint32_t)S.nx * S.ny * S.nz;
    const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ncell) return;
    const int cx = (int)(c % S.nx), cy = (int)((c / S.nx) % S.ny), cz = (int)(c / ((uint32_t)S.nx * S.ny));
    const float lo[3] = {S.ox + cx * S.cell, S.oy + cy * S.cell, S.oz + cz * S.cell};
    const float bo[3] = {S.box_, S.boy_, S.boz_};
    const float sc[3] = {S.bscale_x, S.bscale_y, S.bscale_z};
    const int nd[3] = {S.bnx, S.bny, S.bnz};
    int i0[3], i1[3];
    for (int a = 0; a < 3; ++a) {
        float t0 = (lo[a] - bo[a]) * sc[a] - 0.5f, t1 = (lo[a] + S.cell - bo[a]) * sc[a] - 0.5f;
        i0[a] = max((int)floorf(t0) - 2, 0);
        i1[a] = min((int)floorf(t1) + 3, nd[a] - 1);
    }
    const float4* g = grid4 + S.grid_offset;
    bool far = true;
    for (int k = i0[2]; k <= i1[2] && far; ++k)
        for (int j = i0[1]; j <= i1[1] && far; ++j)
            for (int i = i0[0]; i <= i1[0]; ++i) {
                float4 v = g[(size_t)i + (size_t)nd[0] * ((size_t)j + (size_t)nd[1] * k)];
                if (!(v.x == -S.r0 && v.y == 0.0f && v.z == 0.0f)) {
                    far = false;
                    break;
                }
            }
    mode[S.cell_offset + c] = far ? kWallNone : near_value;
}

}  // namespace

void launch_wall_modes(hipStream_t st, const DevScene& S, const float4* grid4, uint8_t near_value, uint8_t* mode) {
    const uint32_t ncell = (uint32_t)S.nx * S.ny * S.nz;
    hipLaunchKernelGGL(k_wall_modes, dim3((ncell + 255) / 256), dim3(256), 0, st, S, grid4, near_value, mode);
}

void launch_hash(hipStream_t st, const StepBuffers& b) {
    hipLaunchKernelGGL(k_hash, dim3(blocks_for(b.n, kBlock)), dim3(kBlock), 0, st, b.pos_a, b.key, b.n, b.scenes,
                       b.begins, b.nscenes, b.status);
}

void launch_cells_clear(hipStream_t st, const StepBuffers& b) {
    hipLaunchKernelGGL(k_cells_clear, dim3(blocks_for(b.n, kBlock)), dim3(kBlock), 0, st, b.skey, b.n, b.range);
}

void launch_cells(hipStream_t st, const StepBuffers& b) {
    hipLaunchKernelGGL(k_cells, dim3(blocks_for(b.n, kBlock)), dim3(kBlock), 0, st, b.skey, b.n, b.range);
}

void launch_reorder(hipStream_t st, const StepBuffers& b) {
    if (b.nscenes > 1)
        hipLaunchKernelGGL(k_reorder<true>, dim3(blocks_for(b.n, kBlock)), dim3(kBlock), 0, st, b.skey, b.perm,
                           b.range, b.oid_a, b.pos_a, b.vel_a, b.state_a, b.n, b.pos_b, b.vel_b, b.oid_b, b.state_b,
                           b.slot_of_oid);
    else
        hipLaunchKernelGGL(k_reorder<false>, dim3(blocks_for(b.n, kBlock)), dim3(kBlock), 0, st, b.skey, b.perm,
                           b.range, b.oid_a, b.pos_a, b.vel_a, b.state_a, b.n, b.pos_b, b.vel_b, b.oid_b, b.state_b,
                           b.slot_of_oid);
}

// Workgroups of the hit-mask kernels: one thread per slot, or (batch) one
// wavefront per work-list entry.
template <int B, bool ONE>
static unsigned mask_blocks(const StepBuffers& b) {
    return ONE ? blocks_for(b.n, B) : blocks_for(b.ntiles, B / 64);
}

template <int B, bool ONE>
static void launch_density_masks(hipStream_t st, const StepBuffers& b) {
    hipLaunchKernelGGL((k_density_masks<B, ONE>), dim3(mask_blocks<B, ONE>(b)), dim3(B), 0, st, b.pos_b, b.vel_b,
                       b.range, b.n, b.scenes, b.tiles, b.ntiles, b.grid4, b.masks, b.nr, b.split);
}

template <int B, bool ONE>
static void launch_density_one_lane(hipStream_t st, const StepBuffers& b) {
    hipLaunchKernelGGL((k_density<B, ONE>), dim3(blocks_for(b.n, B)), dim3(B), 0, st, b.pos_b, b.vel_b, b.range,
                       b.n, b.scenes, b.begins, b.nscenes, b.grid4, b.nr);
}

void launch_density(hipStream_t st, const StepBuffers& b, int lanes, int block) {
    const bool small = block == 128, one = b.nscenes == 1;
    if (b.masks) {
        if (small) one ? launch_density_masks<128, true>(st, b) : launch_density_masks<128, false>(st, b);
        else one ? launch_density_masks<kBlock, true>(st, b) : launch_density_masks<kBlock, false>(st, b);
        return;
    }
    // the cooperative variants keep 256-thread blocks
    dim3 gc(blocks_for(b.n, kBlock)), tc(kBlock);
    switch (lanes) {
    case 2: hipLaunchKernelGGL(k_density_coop<2>, gc, tc, 0, st, b.pos_b, b.vel_b, b.range, b.n, b.scenes, b.begins, b.nscenes, b.grid4, b.nr); break;
    case 4: hipLaunchKernelGGL(k_density_coop<4>, gc, tc, 0, st, b.pos_b, b.vel_b, b.range, b.n, b.scenes, b.begins, b.nscenes, b.grid4, b.nr); break;
    case 8: hipLaunchKernelGGL(k_density_coop<8>, gc, tc, 0, st, b.pos_b, b.vel_b, b.range, b.n, b.scenes, b.begins, b.nscenes, b.grid4, b.nr); break;
    case 16: hipLaunchKernelGGL(k_density_coop<16>, gc, tc, 0, st, b.pos_b, b.vel_b, b.range, b.n, b.scenes, b.begins, b.nscenes, b.grid4, b.nr); break;
    default:
        if (small) one ? launch_density_one_lane<128, true>(st, b) : launch_density_one_lane<128, false>(st, b);
        else one ? launch_density_one_lane<kBlock, true>(st, b) : launch_density_one_lane<kBlock, false>(st, b);
        break;
    }
}

#define SPH_FORCE_ARGS                                                                                              \
    b.pos_b, b.vel_b, b.range, b.n, b.scenes, b.begins, b.nscenes, b.grid4, b.gridd, b.oid_b, b.state_b, b.acc,    \
        b.pos_a, b.vel_a, b.oid_a, b.state_a, b.key, b.status, b.nr

#define SPH_FORCE_ARGS_M                                                                                            \
    b.pos_b, b.vel_b, b.range, b.n, b.scenes, b.tiles, b.ntiles, b.grid4, b.gridd, b.oid_b, b.state_b, b.acc,      \
        b.pos_a, b.vel_a, b.oid_a, b.state_a, b.key, b.status, b.masks, b.nr

void launch_surface(hipStream_t st, const StepBuffers& b, float cm_thresh, float n_thresh, uint8_t* out) {
    if (b.n == 0) return;
    hipLaunchKernelGGL(k_surface, dim3(blocks_for(b.n, kBlock)), dim3(kBlock), 0, st, b.pos_b, b.range, b.n, b.scenes,
                       b.begins, b.nscenes, b.oid_b, cm_thresh, n_thresh, out);
}

void launch_export_positions(hipStream_t st, uint32_t n, const float4* pos, const float4* vel, const uint32_t* oid,
                             float4* out) {
    if (n == 0) return;
    hipLaunchKernelGGL(k_export_positions, dim3(blocks_for(n, kBlock)), dim3(kBlock), 0, st, n, pos, vel, oid, out);
}

template <int B, bool ONE>
static void launch_force_masks(hipStream_t st, const StepBuffers& b, int cap) {
    dim3 g(mask_blocks<B, ONE>(b)), t(B);
    switch (cap) {
    case 32: hipLaunchKernelGGL((k_force_masks<32, B, ONE>), g, t, 0, st, SPH_FORCE_ARGS_M); break;
    case 64: hipLaunchKernelGGL((k_force_masks<64, B, ONE>), g, t, 0, st, SPH_FORCE_ARGS_M); break;
    case 40: hipLaunchKernelGGL((k_force_masks<40, B, ONE>), g, t, 0, st, SPH_FORCE_ARGS_M); break;
    default: hipLaunchKernelGGL((k_force_masks<48, B, ONE>), g, t, 0, st, SPH_FORCE_ARGS_M); break;
    }
}

void launch_force(hipStream_t st, const StepBuffers& b, int lanes, int block) {
    const bool small = block == 128, one = b.nscenes == 1;
    if (b.masks) {
        // lanes selects the LDS hit-list capacity of the hit-mask variant
        if (small) one ? launch_force_masks<128, true>(st, b, lanes) : launch_force_masks<128, false>(st, b, lanes);
        else one ? launch_force_masks<kBlock, true>(st, b, lanes) : launch_force_masks<kBlock, false>(st, b, lanes);
        return;
    }
    dim3 g(blocks_for(b.n, kBlock)), t(kBlock);
    switch (lanes) {
    case 2: hipLaunchKernelGGL(k_force_coop<2>, g, t, 0, st, SPH_FORCE_ARGS); break;
    case 4: hipLaunchKernelGGL(k_force_coop<4>, g, t, 0, st, SPH_FORCE_ARGS); break;
    case 8: hipLaunchKernelGGL(k_force_coop<8>, g, t, 0, st, SPH_FORCE_ARGS); break;
    case 16: hipLaunchKernelGGL(k_force_coop<16>, g, t, 0, st, SPH_FORCE_ARGS); break;
    default:
        if (small) {
            if (one) hipLaunchKernelGGL((k_force<128, true>), dim3(blocks_for(b.n, 128)), dim3(128), 0, st, SPH_FORCE_ARGS);
            else hipLaunchKernelGGL((k_force<128, false>), dim3(blocks_for(b.n, 128)), dim3(128), 0, st, SPH_FORCE_ARGS);
        } else {
            if (one) hipLaunchKernelGGL((k_force<kBlock, true>), g, t, 0, st, SPH_FORCE_ARGS);
            else hipLaunchKernelGGL((k_force<kBlock, false>), g, t, 0, st, SPH_FORCE_ARGS);
        }
        break;
    }
}

#undef SPH_FORCE_ARGS
#undef SPH_FORCE_ARGS_M

}  // namespace sph
