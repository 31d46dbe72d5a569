// kernels.h -- host-side launchers of the SPH step kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "common.h"

namespace sph {

// Device buffers of one particle set (one or many scenes), cell-sorted.
struct StepBuffers {
    uint32_t n = 0;              // particles (all scenes)
    int nscenes = 0;
    const DevScene* scenes = nullptr;     // [nscenes]
    const uint32_t* begins = nullptr;     // [nscenes] first slot of each scene
    // batch work list of the hit-mask kernels: per wavefront (first slot,
    // scene << 7 | slot count), never across a scene boundary (nscenes > 1):
    // the aligned 64-slot wavefronts, then the scenes' starts inside them
    const uint2* tiles = nullptr;
    uint32_t ntiles = 0;
    const float4* grid4 = nullptr;        // (distance, density, viscosity, 0) pool
    const float* gridd = nullptr;         // distance pool
    // storage order ("a") and cell-sorted order ("b")
    float4* pos_a = nullptr;
    float4* vel_a = nullptr;     // (vx, vy, vz, density of last step)
    uint32_t* oid_a = nullptr;
    uint8_t* state_a = nullptr;
    float4* pos_b = nullptr;
    float4* vel_b = nullptr;     // (vx, vy, vz, density)
    uint32_t* oid_b = nullptr;
    uint8_t* state_b = nullptr;
    float4* acc = nullptr;       // sorted order
    uint32_t* key = nullptr;     // keys of the storage order
    uint32_t* skey = nullptr;    // sorted keys
    uint32_t* perm = nullptr;    // storage slot of each sorted slot (before tie fix-up)
    uint32_t* slot_of_oid = nullptr;
    int2* range = nullptr;       // cell -> [start, end) in sorted order
    uint32_t* status = nullptr;
    // row words of the force neighbours (format in common.h), written by the
    // density pass and read by the force pass; null = neighbour loops
    // without masks
    uint64_t* masks = nullptr;
    // two-scale split flags, set by the hit-mask density pass (state == null: none)
    SplitRefs split = {nullptr, nullptr, nullptr, nullptr};
    // learned boundary records (mode == null: grid everywhere)
    NeuralRefs nr = {nullptr, nullptr, nullptr, nullptr, nullptr, 0};
};

void launch_hash(hipStream_t st, const StepBuffers& b);
void launch_cells_clear(hipStream_t st, const StepBuffers& b);
void launch_cells(hipStream_t st, const StepBuffers& b);
void launch_reorder(hipStream_t st, const StepBuffers& b);
// Boundary source of every cell of scene S (mode + S.cell_offset ...): none
// where the cell's reachable texels hold only the far-field values, else
// near_value.
void launch_wall_modes(hipStream_t st, const DevScene& S, const float4* grid4, uint8_t near_value, uint8_t* mode);
// lanes: lanes per particle of the neighbour loops (1 = one thread per
// particle; 2, 4, 8, 16 = wavefront-cooperative variants).  block: threads
// per workgroup of the one-lane and hit-mask variants (128 or 256; the
// cooperative variants keep 256).  With b.masks set
// the hit-mask pair is launched instead and the force's lanes value selects
// its per-thread hit-list capacity (32, 48 or 64).
void launch_density(hipStream_t st, const StepBuffers& b, int lanes, int block = 256);
// Surface flag per original id from the sorted positions and cell ranges of
// the last step (k_surface).
void launch_surface(hipStream_t st, const StepBuffers& b, float cm_thresh, float n_thresh, uint8_t* out);
// out[oid[i]] = (pos[i].xyz, vel[i].w) for i < n: positions by original id.
void launch_export_positions(hipStream_t st, uint32_t n, const float4* pos, const float4* vel, const uint32_t* oid,
                             float4* out);
void launch_force(hipStream_t st, const StepBuffers& b, int lanes, int block = 256);

}  // namespace sph
