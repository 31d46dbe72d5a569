// kernels.hip -- the coarse SPH step on gfx950.
//
// Data layout: particles live in SoA float4 streams that are re-ordered by
// cell every step, so that neighbour gathers are coalesced and every
// 3-cell x-row of a neighbourhood is one contiguous run of slots.
//   pos (x, y, z, 0), vel (vx, vy, vz, density), acc (ax, ay, az, 0).
// The reference keeps AoS arrays in original-id order and gathers through
// the sorted id list instead (particle_simulation.cu:206,327).
//
// Per step:  [hash, first step only] -> clear last step's cells ->
//   radix sort (key, storage slot) -> cell ranges -> reorder (+ tie fix-up
//   to original-id order) -> density -> force + integrate + collision +
//   next key.
#include "kernels.h"

namespace sph {

namespace {

constexpr int kBlock = 256;

inline unsigned blocks_for(uint32_t n, int block) { return (n + block - 1) / block; }

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

// Gather storage -> sorted order.  The radix sort is stable with respect to
// the storage order; the reference sorts (hash, original id), so inside
// each cell run the slot is the rank of the original id (runs are short:
// a cell holds ~8 particles at rest density).
__global__ __launch_bounds__(kBlock) void k_reorder(
    const uint32_t* __restrict__ skey, const uint32_t* __restrict__ perm, const int2* __restrict__ range,
    const uint32_t* __restrict__ oid_a, const float4* __restrict__ pos_a, const float4* __restrict__ vel_a,
    const uint8_t* __restrict__ state_a, uint32_t n, float4* __restrict__ pos_b, float4* __restrict__ vel_b,
    uint32_t* __restrict__ oid_b, uint8_t* __restrict__ state_b, uint32_t* __restrict__ slot_of_oid) {
    uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    uint32_t k = skey[s];
    int2 r = range[k];
    uint32_t src = perm[s];
    uint32_t me = oid_a[src];
    uint32_t rank = 0;
    for (int q = r.x; q < r.y; ++q) rank += (oid_a[perm[q]] < me) ? 1u : 0u;
    uint32_t dst = (uint32_t)r.x + rank;
    pos_b[dst] = pos_a[src];
    vel_b[dst] = vel_a[src];
    oid_b[dst] = me;
    if (state_a) state_b[dst] = state_a[src];
    if (slot_of_oid) slot_of_oid[me] = dst;
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

// compute_particle_density_pressure (cu:196-256, 1385-1419).  Density goes
// to vel.w; pressure k(rho - rho0) is recomputed where needed.
__global__ __launch_bounds__(kBlock) void k_density(const float4* __restrict__ pos, float4* __restrict__ vel,
                                                    const int2* __restrict__ range, uint32_t n,
                                                    const DevScene* __restrict__ scenes,
                                                    const uint32_t* __restrict__ begins, int nscenes,
                                                    const float4* __restrict__ grid4) {
    uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    const DevScene& S = scenes[scene_of(begins, nscenes, s)];
    float4 xi = pos[s];
    const float h = S.h, h2 = S.h2;
    int lx = range_coord(xi.x, -h, S.ox, S.cell, S.nx), hx = range_coord(xi.x, h, S.ox, S.cell, S.nx);
    int ly = range_coord(xi.y, -h, S.oy, S.cell, S.ny), hy = range_coord(xi.y, h, S.oy, S.cell, S.ny);
    int lz = range_coord(xi.z, -h, S.oz, S.cell, S.nz), hz = range_coord(xi.z, h, S.oz, S.cell, S.nz);
    float sum = 0.0f;
    for (int k = lz; k <= hz; ++k) {
        for (int j = ly; j <= hy; ++j) {
            uint32_t rowbase = S.cell_offset + (uint32_t)(S.nx * (S.ny * k + j));
            int b, e;
            row_run(range, rowbase, lx, hx, b, e);
            for (int q = b; q < e; ++q) {
                float4 xj = pos[q];
                float dx = xj.x - xi.x, dy = xj.y - xi.y, dz = xj.z - xi.z;
                float r2 = dx * dx + dy * dy + dz * dz;
                float d = fmaxf(h2 - r2, 0.0f);
                sum += d * d * d;
            }
        }
    }
    float rho = sum * (S.poly6 * S.mass);
    float tx, ty, tz;
    texel_coords(S, xi.x, xi.y, xi.z, tx, ty, tz);
    float4 g = lerp3<float4>(grid4, S, tex_setup(S, tx, ty, tz));
    rho += g.y;
    reinterpret_cast<float*>(vel)[4 * (size_t)s + 3] = rho;
}

// compute_particle_acceleration_ifsurf (cu:315-474, 1459-1556) fused with
// integrate_euler (cu:632-665), the collision response (collision_handling,
// cu:742-834, driven by the distance grid) and the next step's cell key.
// Writes the next storage order (= this step's sorted order).
__global__ __launch_bounds__(kBlock) void k_force(
    const float4* __restrict__ pos, const float4* __restrict__ vel, const int2* __restrict__ range, uint32_t n,
    const DevScene* __restrict__ scenes, const uint32_t* __restrict__ begins, int nscenes,
    const float4* __restrict__ grid4, const float* __restrict__ gridd, const uint32_t* __restrict__ oid_b,
    const uint8_t* __restrict__ state_b, float4* __restrict__ acc, float4* __restrict__ pos_out,
    float4* __restrict__ vel_out, uint32_t* __restrict__ oid_out, uint8_t* __restrict__ state_out,
    uint32_t* __restrict__ key_out, uint32_t* __restrict__ status) {
    uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    const DevScene& S = scenes[scene_of(begins, nscenes, s)];
    const float4 xi = pos[s];
    const float4 vi = vel[s];
    const float rhoi = vi.w;
    const float h = S.h, h2 = S.h2;
    const float Ai = S.stiff * (rhoi - S.rho0) / rhoi;   // p_i / rho_i
    const float kr0 = S.stiff * S.rho0;
    int lx = range_coord(xi.x, -h, S.ox, S.cell, S.nx), hx = range_coord(xi.x, h, S.ox, S.cell, S.nx);
    int ly = range_coord(xi.y, -h, S.oy, S.cell, S.ny), hy = range_coord(xi.y, h, S.oy, S.cell, S.ny);
    int lz = range_coord(xi.z, -h, S.oz, S.cell, S.nz), hz = range_coord(xi.z, h, S.oz, S.cell, S.nz);
    float fpx = 0, fpy = 0, fpz = 0;   // sum (A_i + rho_i p_j/rho_j^2) (h-r)^2 / r * r_ij
    float fvx = 0, fvy = 0, fvz = 0;   // sum (h-r)/rho_j * (v_j - v_i)
    float cgx = 0, cgy = 0, cgz = 0;   // sum (h^2-r^2)^2/rho_j * r_ij
    float cl = 0;                      // sum (h^2-r^2)(3h^2-7r^2)/rho_j
    for (int k = lz; k <= hz; ++k) {
        for (int j = ly; j <= hy; ++j) {
            uint32_t rowbase = S.cell_offset + (uint32_t)(S.nx * (S.ny * k + j));
            int b, e;
            row_run(range, rowbase, lx, hx, b, e);
            for (int q = b; q < e; ++q) {
                float4 xj = pos[q];
                float dx = xi.x - xj.x, dy = xi.y - xj.y, dz = xi.z - xj.z;
                float r2 = dx * dx + dy * dy + dz * dz;
                if (r2 <= h2 && r2 > 0.0f) {
                    float4 vj = vel[q];
                    float inv_r = __builtin_amdgcn_rsqf(r2);
                    float rn = r2 * inv_r;
                    float invj = __builtin_amdgcn_rcpf(vj.w);
                    float pj = S.stiff * vj.w - kr0;
                    float hr = h - rn;
                    float qd = h2 - r2;
                    float cp = (Ai + rhoi * pj * invj * invj) * hr * hr * inv_r;
                    fpx += cp * dx; fpy += cp * dy; fpz += cp * dz;
                    float wv = invj * hr;
                    fvx += wv * (vj.x - vi.x); fvy += wv * (vj.y - vi.y); fvz += wv * (vj.z - vi.z);
                    float wc = invj * qd * qd;
                    cgx += wc * dx; cgy += wc * dy; cgz += wc * dz;
                    cl += invj * qd * (3.0f * h2 - 7.0f * r2);
                }
            }
        }
    }
    const float cpf = -S.mass * S.grad_spiky, cvf = S.mu * S.mass * S.lapl_visc;
    float Fx = cvf * fvx + cpf * fpx, Fy = cvf * fvy + cpf * fpy, Fz = cvf * fvz + cpf * fpz;
    float tx, ty, tz;
    texel_coords(S, xi.x, xi.y, xi.z, tx, ty, tz);
    float4 g = lerp3<float4>(grid4, S, tex_setup(S, tx, ty, tz));
    float distWall = -g.x;
    if (distWall < S.r0) {
        // wall penalty and wall viscosity (cu:407-452)
        float coeff = rhoi / (S.dt * S.dt) * (S.r0 - distWall);
        float3 gn = dist_grad_normalized(gridd, S, tx, ty, tz);
        Fx -= coeff * gn.x + vi.x * g.z;
        Fy -= coeff * gn.y + vi.y * g.z;
        Fz -= coeff * gn.z + vi.z * g.z;
    } else {
        // surface tension (cu:454-466)
        float cm = S.mass * S.grad_poly6;
        float gx = cm * cgx, gy = cm * cgy, gz = cm * cgz;
        float gnorm = sqrtf(gx * gx + gy * gy + gz * gz);
        if (gnorm > S.norm_thresh) {
            float fc = S.sigma * (S.mass * S.lapl_poly6 * cl) / gnorm;
            Fx -= fc * gx;
            Fy -= fc * gy;
            Fz -= fc * gz;
        }
    }
    float ax = Fx / rhoi, ay = Fy / rhoi - S.gravity, az = Fz / rhoi;
    acc[s] = make_float4(ax, ay, az, 0.0f);
    // integrate_euler: v += dt a; x += dt v
    float vx = vi.x + S.dt * ax, vy = vi.y + S.dt * ay, vz = vi.z + S.dt * az;
    float px = xi.x + S.dt * vx, py = xi.y + S.dt * vy, pz = xi.z + S.dt * vz;
    if (S.collision) {
        float cx, cy, cz;
        texel_coords(S, px, py, pz, cx, cy, cz);
        float D = dist_at(gridd, S, cx, cy, cz);
        if (D > 0.0f) {
            float3 gw = dist_grad_normalized(gridd, S, cx, cy, cz);
            // normalized-coordinate gradient -> world gradient: divide by L
            gw.x *= S.bscale_x * (1.0f / (float)S.bnx);
            gw.y *= S.bscale_y * (1.0f / (float)S.bny);
            gw.z *= S.bscale_z * (1.0f / (float)S.bnz);
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
    pos_out[s] = make_float4(px, py, pz, 0.0f);
    vel_out[s] = make_float4(vx, vy, vz, rhoi);
    key_out[s] = kk;
    oid_out[s] = oid_b[s];
    if (state_out) state_out[s] = state_b[s];
}

}  // namespace

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
    hipLaunchKernelGGL(k_reorder, dim3(blocks_for(b.n, kBlock)), dim3(kBlock), 0, st, b.skey, b.perm, b.range,
                       b.oid_a, b.pos_a, b.vel_a, b.state_a, b.n, b.pos_b, b.vel_b, b.oid_b, b.state_b,
                       b.slot_of_oid);
}

void launch_density(hipStream_t st, const StepBuffers& b, int block) {
    (void)block;
    hipLaunchKernelGGL(k_density, dim3(blocks_for(b.n, kBlock)), dim3(kBlock), 0, st, b.pos_b, b.vel_b, b.range,
                       b.n, b.scenes, b.begins, b.nscenes, b.grid4);
}

void launch_force(hipStream_t st, const StepBuffers& b, int block) {
    (void)block;
    hipLaunchKernelGGL(k_force, dim3(blocks_for(b.n, kBlock)), dim3(kBlock), 0, st, b.pos_b, b.vel_b, b.range,
                       b.n, b.scenes, b.begins, b.nscenes, b.grid4, b.gridd, b.oid_b, b.state_b, b.acc, b.pos_a,
                       b.vel_a, b.oid_a, b.state_a, b.key, b.status);
}

}  // namespace sph
