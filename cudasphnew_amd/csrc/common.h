// common.h -- device-side types and helpers shared by the SPH kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

#include "../../include/sph.h"

namespace sph {

constexpr uint32_t kEmpty = SPH_EMPTY_CELL;
constexpr uint32_t kStatusOutOfDomain = 1u;
constexpr uint32_t kStatusNonFinite = 2u;
constexpr uint32_t kStatusCapacity = 4u;

// Per-scene constants, one record per scene in device memory.  Replaces the
// module-global __constant__ SimulationParameters and boundary constants of
// the reference (particle_simulation.cu:26,42-46) so that any number of
// scenes (and handles) can coexist.
struct DevScene {
    // cell grid (gridOrigin/gridSpacing/gridDim, h:56-58)
    float ox, oy, oz, cell;
    int nx, ny, nz;
    uint32_t cell_offset;      // first global key of this scene
    // sub-particle cell grid (h:59-60)
    float cell_sub;
    int snx, sny, snz;
    // physics (h:64-98)
    float h, h2, mass, stiff, rho0, mu, gravity, sigma, norm_thresh, dt;
    float poly6, grad_poly6, lapl_poly6, grad_spiky, lapl_visc;
    float h_sub, h2_sub, mass_sub, dt_sub;
    float poly6_sub, grad_poly6_sub, lapl_poly6_sub, grad_spiky_sub, lapl_visc_sub;
    // boundary grid (cu:42-53): texel coordinate t = (x - bo) * bscale - 0.5
    float box_, boy_, boz_;
    float bscale_x, bscale_y, bscale_z;  // dims / L
    float gstep_x, gstep_y, gstep_z;     // dims / (dims - 1): texels per spacing
    float ghalf_x, ghalf_y, ghalf_z;     // (dims - 1) / 2
    float bspacing, r0;
    int bnx, bny, bnz;
    uint32_t dist_shift;                  // distance of node i at gridd[i << dist_shift]: 0 for the
                                          // float copy, 2 when gridd is the float4 grid (channel x)
    uint64_t grid_offset;                 // first node in the grid pool
    float restitution;
    int collision;
    // two-scale split region
    float split_lo[3], split_hi[3];
    int sub_placement;
    uint32_t p_begin, p_end;              // particle range of the scene
};

// Raw buffer view of a float4 array: a gather takes a 32-bit byte offset
// (one shift, and one offset register shared by the position and velocity
// gathers of a neighbour) instead of a 64-bit address per load.  Slots stay
// below 2^28 (the engine's particle and sub-particle limits).  Measured at 1M particles: force
// pass 95 → 91 us.
struct F4Buf {
    __amdgpu_buffer_rsrc_t r;
    __device__ __forceinline__ explicit F4Buf(const float4* p)
        : r(__builtin_amdgcn_make_buffer_rsrc(const_cast<float4*>(p), (short)0, -1, 0x00020000)) {}
    __device__ __forceinline__ float4 operator[](uint32_t q) const {
        return __builtin_bit_cast(float4, __builtin_amdgcn_raw_buffer_load_b128(r, (int)(q * 16u), 0, 0));
    }
};

// Row words of the hit-mask pair, written by the density pass for each
// particle and each of its nine neighbour rows (a run of consecutive slots),
// read by the force pass (and the two-scale marking pass).  Buffer layout:
// [9][n] 64-bit words, then [3][9][n] 32-bit chunk masks.
//   word = first slot << 32 | mask of candidates 0..31 | (chunks - 1) << 60
// where bit t of a mask is set when candidate t of the run is within h (own
// slot excluded), and a row of c candidates has ceil(c / 32) chunks: chunk
// k >= 1 (candidates 32k..32k+31) sits at chunk_masks(masks, n)[chunk_at(n,
// k, row, slot)].  Near rest density a row holds at most 32 candidates (24
// on a lattice) and only the word is written; a compressed flow's rows reach
// 40-90.  A row of more than kRowMaxCands candidates gets kRowLong and is
// scanned directly.  First slots are < 2^28, so bits 62..63 of a word are
// clear and no word equals kRowLong.
constexpr int kRowMaxChunks = 4;
constexpr int kRowMaxCands = 32 * kRowMaxChunks;
constexpr uint64_t kRowLong = ~0ull;
// allocation of the buffer in 64-bit units
constexpr size_t row_words_u64(size_t n) { return 9 * n + ((kRowMaxChunks - 1) * 9 * n + 1) / 2; }

__device__ __forceinline__ uint32_t row_first(uint64_t w) { return (uint32_t)(w >> 32) & 0x0FFFFFFFu; }
__device__ __forceinline__ uint32_t row_last_chunk(uint64_t w) { return (uint32_t)(w >> 60) & 3u; }
template <class T>
__device__ __forceinline__ auto chunk_masks(T* masks, uint32_t n) {
    using U = typename std::conditional<std::is_const<T>::value, const uint32_t, uint32_t>::type;
    return reinterpret_cast<U*>(masks + 9 * (size_t)n);
}
__device__ __forceinline__ size_t chunk_at(uint32_t n, uint32_t k, int row, uint32_t s) {
    return ((size_t)(k - 1) * 9 + (size_t)row) * n + s;
}

// Boundary source of a cell when a learned SDF is attached: no wall term (far
// from every surface), the network's per-particle record, or the grid.
enum : uint8_t { kWallNone = 0, kWallNeural = 1, kWallGrid = 2 };

// The two-scale state pass fused into the hit-mask density pass
// (shift_state + find_split_particles, particle_simulation.cu:730-740,
// 932-1001) on this step's sorted particles; state == null: no two-scale
// step.
struct SplitRefs {
    uint8_t* state;        // per sorted slot: state << 2, split bit set here
    uint8_t* state_fin;    // the final state's starting value (the boundary bit comes later)
    uint32_t* nmark;       // length of the split list (zero before the pass)
    uint32_t* mark_list;   // slots of the split particles, one atomic per wavefront
};

struct NeuralRefs {
    const uint8_t* mode;    // per global cell key; null: grid everywhere
    const uint32_t* skey;   // sorted keys
    const float4* rec;      // per sorted slot: central difference of the learned distance at
                            // +-spacing (world units, the grid normal's stencil), 0
    uint8_t* need;          // per sorted slot: learned-source particle in the wall-penalty
                            // zone (written by the density pass, read by the network pass)
    uint32_t* list_count;   // per scene: length of the network pass's particle list,
                            // zeroed by the density pass that flags the particles
    int nscenes;
};

__device__ __forceinline__ int wall_mode(const NeuralRefs& nr, uint32_t s) {
    return nr.mode ? (int)nr.mode[nr.skey[s]] : (int)kWallGrid;
}

// ---------------------------------------------------------------------------
// Cell coordinates.  These must produce bit-identical keys to the FP32
// reference expressions; no fast-math, IEEE division.

// compute_particle_hash per-axis coordinate (particle_simulation.cu:101-115):
// (int)((x - o) / d) % dim.  A coordinate <= -1 (negative hash in the
// reference) becomes 0, >= dim wraps like "% dim"; both flag out-of-domain.
__device__ __forceinline__ int hash_coord(float x, float o, float d, int dim, uint32_t& flags) {
    float f = (x - o) / d;
    if (!(f == f)) { flags |= kStatusNonFinite; return 0; }
    if (f <= -1.0f) { flags |= kStatusOutOfDomain; return 0; }
    if (f >= 2147483520.0f) { flags |= kStatusOutOfDomain; return 0; }
    int c = (int)f;
    if (c >= dim) { flags |= kStatusOutOfDomain; c %= dim; }
    return c;
}

// compute_grid_coordinate (particle_simulation.cu:1238-1258):
// (unsigned)((x + d - o) / s) % dim, clamped; the float->unsigned
// conversion saturates at 0 for negative and NaN values.
__device__ __forceinline__ int range_coord(float x, float dd, float o, float s, int dim) {
    float f = ((x + dd) - o) / s;
    if (!(f > 0.0f)) return 0;
    if (f >= 2147483520.0f) return 0;
    int c = (int)((unsigned)f) % dim;
    return min(max(c, 0), dim - 1);
}

__device__ __forceinline__ uint32_t cell_key(const DevScene& S, float x, float y, float z, uint32_t& flags) {
    int cx = hash_coord(x, S.ox, S.cell, S.nx, flags);
    int cy = hash_coord(y, S.oy, S.cell, S.ny, flags);
    int cz = hash_coord(z, S.oz, S.cell, S.nz, flags);
    return S.cell_offset + (uint32_t)(S.nx * (S.ny * cz + cy) + cx);
}

__device__ __forceinline__ uint32_t sub_cell_key(const DevScene& S, float x, float y, float z,
                                                 uint32_t& flags) {
    int cx = hash_coord(x, S.ox, S.cell_sub, S.snx, flags);
    int cy = hash_coord(y, S.oy, S.cell_sub, S.sny, flags);
    int cz = hash_coord(z, S.oz, S.cell_sub, S.snz, flags);
    return (uint32_t)(S.snx * (S.sny * cz + cy) + cx);
}

// XCD-aware block order.  The 8 XCDs of an MI355X receive workgroups
// round-robin (block b runs on XCD b % 8), each XCD with its own L2.  Giving
// XCD x the contiguous logical blocks [x*G/8, (x+1)*G/8) keeps neighbouring
// particle tiles -- which share their neighbour rows -- in one L2.
__device__ __forceinline__ uint32_t xcd_block(uint32_t b, uint32_t G) {
    constexpr uint32_t X = 8;
    uint32_t per = G / X, rem = G % X;
    uint32_t x = b % X, i = b / X;
    return x * per + min(x, rem) + i;
}

// Scene of a particle slot: scenes occupy fixed contiguous slot ranges in
// both storage and sorted order (their keys are disjoint, increasing ranges).
__device__ __forceinline__ int scene_of(const uint32_t* __restrict__ begins, int nscenes, uint32_t s) {
    int lo = 0, hi = nscenes - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (begins[mid] <= s) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// ---------------------------------------------------------------------------
// Boundary grid lookup: normalized coordinates, linear filter, clamp
// addressing (set_up_3d_float_texture, particle_simulation.cu:3003-3007),
// in software with exact FP32 weights.  tx,ty,tz are texel coordinates
// (already u*dims - 0.5).

struct TexCoord {
    int x0, x1, y0, y1, z0, z1;
    float ax, ay, az;
};

__device__ __forceinline__ void tex_setup1(float t, int n, int& i0, int& i1, float& a) {
    t = fminf(fmaxf(t, -1.0f), (float)n);
    float fl = floorf(t);
    a = t - fl;
    int b = (int)fl;
    i0 = min(max(b, 0), n - 1);
    i1 = min(max(b + 1, 0), n - 1);
}

__device__ __forceinline__ TexCoord tex_setup(const DevScene& S, float tx, float ty, float tz) {
    TexCoord c;
    tex_setup1(tx, S.bnx, c.x0, c.x1, c.ax);
    tex_setup1(ty, S.bny, c.y0, c.y1, c.ay);
    tex_setup1(tz, S.bnz, c.z0, c.z1, c.az);
    return c;
}

// a + t (b - a): reproduces a constant field exactly, so the clamped
// "far from any wall" value -r0 interpolates to exactly -r0 and the
// wall-term switch distWall < r0 (particle_simulation.cu:422) is not flipped
// by rounding.
__device__ __forceinline__ float lerp1(float a, float b, float t) { return a + t * (b - a); }

template <class T>
__device__ __forceinline__ T lerp3(const T* __restrict__ g, const DevScene& S, const TexCoord& c);

// The distance grid: its own float copy, or channel x of the float4 grid
// (S.dist_shift = 2) once a learned boundary has replaced the copy.
template <>
__device__ __forceinline__ float lerp3<float>(const float* __restrict__ g, const DevScene& S,
                                              const TexCoord& c) {
    const uint32_t sh = S.dist_shift;
    size_t sx = (size_t)1 << sh, sy = (size_t)S.bnx << sh, sz = ((size_t)S.bnx * S.bny) << sh;
    const float* b = g + (S.grid_offset << sh);
    float v000 = b[c.x0 * sx + c.y0 * sy + c.z0 * sz], v100 = b[c.x1 * sx + c.y0 * sy + c.z0 * sz];
    float v010 = b[c.x0 * sx + c.y1 * sy + c.z0 * sz], v110 = b[c.x1 * sx + c.y1 * sy + c.z0 * sz];
    float v001 = b[c.x0 * sx + c.y0 * sy + c.z1 * sz], v101 = b[c.x1 * sx + c.y0 * sy + c.z1 * sz];
    float v011 = b[c.x0 * sx + c.y1 * sy + c.z1 * sz], v111 = b[c.x1 * sx + c.y1 * sy + c.z1 * sz];
    float c00 = lerp1(v000, v100, c.ax), c10 = lerp1(v010, v110, c.ax);
    float c01 = lerp1(v001, v101, c.ax), c11 = lerp1(v011, v111, c.ax);
    return lerp1(lerp1(c00, c10, c.ay), lerp1(c01, c11, c.ay), c.az);
}

__device__ __forceinline__ float4 lerp4(float4 a, float4 b, float t) {
    return make_float4(lerp1(a.x, b.x, t), lerp1(a.y, b.y, t), lerp1(a.z, b.z, t), lerp1(a.w, b.w, t));
}

template <>
__device__ __forceinline__ float4 lerp3<float4>(const float4* __restrict__ g, const DevScene& S,
                                                const TexCoord& c) {
    size_t sx = 1, sy = (size_t)S.bnx, sz = (size_t)S.bnx * S.bny;
    const float4* b = g + S.grid_offset;
    float4 c00 = lerp4(b[c.x0 * sx + c.y0 * sy + c.z0 * sz], b[c.x1 * sx + c.y0 * sy + c.z0 * sz], c.ax);
    float4 c10 = lerp4(b[c.x0 * sx + c.y1 * sy + c.z0 * sz], b[c.x1 * sx + c.y1 * sy + c.z0 * sz], c.ax);
    float4 c01 = lerp4(b[c.x0 * sx + c.y0 * sy + c.z1 * sz], b[c.x1 * sx + c.y0 * sy + c.z1 * sz], c.ax);
    float4 c11 = lerp4(b[c.x0 * sx + c.y1 * sy + c.z1 * sz], b[c.x1 * sx + c.y1 * sy + c.z1 * sz], c.ax);
    return lerp4(lerp4(c00, c10, c.ay), lerp4(c01, c11, c.ay), c.az);
}

__device__ __forceinline__ void texel_coords(const DevScene& S, float x, float y, float z, float& tx,
                                             float& ty, float& tz) {
    tx = (x - S.box_) * S.bscale_x - 0.5f;
    ty = (y - S.boy_) * S.bscale_y - 0.5f;
    tz = (z - S.boz_) * S.bscale_z - 0.5f;
}

// Distance at a texel coordinate.
__device__ __forceinline__ float dist_at(const float* __restrict__ gd, const DevScene& S, float tx, float ty,
                                         float tz) {
    return lerp3<float>(gd, S, tex_setup(S, tx, ty, tz));
}

// Central-difference gradient of the distance grid in normalized texture
// coordinates (particle_simulation.cu:429-439): the lookups sit
// +-spacing/L away in u, i.e. +-dims/(dims-1) texels; the quotient by 2*dX
// is (dims-1)/2 per texel difference.
__device__ __forceinline__ float3 dist_grad_normalized(const float* __restrict__ gd, const DevScene& S,
                                                       float tx, float ty, float tz) {
    float3 g;
    g.x = (dist_at(gd, S, tx + S.gstep_x, ty, tz) - dist_at(gd, S, tx - S.gstep_x, ty, tz)) * S.ghalf_x;
    g.y = (dist_at(gd, S, tx, ty + S.gstep_y, tz) - dist_at(gd, S, tx, ty - S.gstep_y, tz)) * S.ghalf_y;
    g.z = (dist_at(gd, S, tx, ty, tz + S.gstep_z) - dist_at(gd, S, tx, ty, tz - S.gstep_z)) * S.ghalf_z;
    return g;
}

}  // namespace sph
