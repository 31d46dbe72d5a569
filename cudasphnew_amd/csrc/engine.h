// engine.h -- host orchestration of the SPH step for one or many scenes.
// Replaces ParticleSimulation's Init/Advance/stage wrappers
// (particle_simulation.cu:1882-2117, 2320-2590).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

#include "common.h"
#include "cell_sort.h"
#include "kernels.h"
#include "radix_sort.h"

namespace sph {

enum Stage {
    kStageHash = 0,
    kStageClear,
    kStageSort,
    kStageCells,
    kStageReorder,
    kStageDensity,
    kStageForce,
    kStageTwoScaleState,
    kStageTwoScaleCollect,
    kStageTwoScaleSeed,
    kStageTwoScaleSubSort,
    kStageTwoScaleSubPhysics,
    kStageWallMlp,
    kStageSubStepProject,   // sub-particle step: state back to list order + projection
    kNumStages
};

// Timing marks that open an interval: a step (counted in the per-step
// averages) or a sub-particle step (its stages add to the same totals but it
// is not counted as a step).
constexpr int kMarkStep = -1;
constexpr int kMarkSubStep = -2;

const char* stage_name(int s);

struct TwoScale;  // two-scale pass state (twoscale.hip)
struct NeuralBoundary;  // learned boundary state (neural.hip)

struct Engine {
    int device = 0;
    hipStream_t stream = nullptr;
    int nscenes = 0;
    std::vector<sph_params> params;
    std::vector<DevScene> hscenes;
    std::vector<uint32_t> hbegins;   // nscenes + 1
    std::vector<size_t> cells_of;    // per scene
    uint32_t N = 0;
    uint32_t total_cells = 0;
    int key_bits = 0;

    // device memory
    DevScene* d_scenes = nullptr;
    uint32_t* d_begins = nullptr;
    uint2* d_tiles = nullptr;            // batch work list of the hit-mask kernels (nscenes > 1)
    uint32_t ntiles = 0;
    float4* grid4 = nullptr;
    float* gridd = nullptr;                // float distance copy, or grid4 itself when gridd_alias
    bool gridd_alias = false;
    size_t grid_nodes = 0;
    float4 *pos_a = nullptr, *vel_a = nullptr, *pos_b = nullptr, *vel_b = nullptr, *acc = nullptr;
    uint32_t *oid_a = nullptr, *oid_b = nullptr, *key = nullptr, *skey = nullptr, *perm = nullptr;
    uint32_t* slot_of_oid = nullptr;
    uint8_t *state_a = nullptr, *state_b = nullptr;
    int2* range = nullptr;
    uint8_t* wall_mode = nullptr;   // per global cell: kWallNone / kWallNeural / kWallGrid
    uint32_t* status = nullptr;
    uint32_t* n_dev = nullptr;
    RadixSort sorter;
    // counting sort over the cell table, used when the key space is at most
    // kCellSortRatio cells per particle (every workload here: 2.5-9); the
    // radix sort otherwise
    CellSort csort;
    bool use_cell_sort = false;
    size_t bytes = 0;

    bool keys_valid = false;
    bool cells_valid = false;
    uint32_t steps = 0;
    std::string err;

    // timing: events recorded on the stream before each step and after
    // each stage (stage -1 marks a step start); resolved without per-step
    // host synchronisation when the times are requested.
    bool timing = false;
    std::vector<hipEvent_t> ev_pool;
    size_t ev_used = 0;
    std::vector<std::pair<int, hipEvent_t>> ev_rec;
    float stage_acc_ms[kNumStages] = {};
    uint32_t stage_launches[kNumStages] = {};
    uint32_t stage_steps = 0;
    hipEvent_t next_event();
    sph_status resolve_timing();

    // graph capture of the steady-state step, plain and two-scale
    bool use_graph = false;
    uint32_t retune_every = 0;   // re-run the autotuner before every n-th step (0: never)
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;

    TwoScale* ts = nullptr;
    bool last_two_scale = false;
    NeuralBoundary* nb = nullptr;

    // kernel variants of the density / force passes
    sph_tuning tuning = {1, 1, 0, 48, 256};
    uint64_t* masks = nullptr;   // row words [9][N] + chunk masks [3][9][N] (common.h)

    // render interop buffers, allocated on the first map
    float4* render_pos = nullptr;   // [N] by original id
    float4* render_sub = nullptr;   // [8N] by sub-particle id
    uint8_t* surface = nullptr;     // [N] surface flags by original id, on demand

    ~Engine();
    void drop_graph();
    sph_status upload_scenes();
    sph_status alias_distance();   // release the float distance copy (every scene has a network)
    sph_status own_distance();     // allocate and fill it again from grid4
    sph_status set_tuning(const sph_tuning& t);
    sph_status autotune(int reps, sph_tuning* out);
    sph_status init(int S, const sph_params* P, const sph_scene* scenes);
    sph_status step(int n, bool two_scale);
    sph_status sync();
    sph_status check_launch(const char* what);
    StepBuffers buffers() const;
    void enqueue_step(bool two_scale);
    void mark(int stage);
    sph_status fail(sph_status s, const std::string& msg);
    template <class T>
    sph_status alloc(T** p, size_t count);
    sph_status read_field(int scene, sph_field f, void* dst, size_t bytes_out);
    sph_status upload_params(int scene, const sph_params& p);
    sph_status map_render(sph_render_view* out);
    sph_status extract_surface(uint8_t* flags, uint32_t* count);
};

DevScene make_dev_scene(const sph_params& P, const sph_boundary& B);
// The message sph_last_error_message(NULL) returns on this thread (errors of
// calls without a handle: sph_create, sph_step_noalloc).
void set_global_error(const std::string& msg);

// Two-scale pass (twoscale.hip).
sph_status twoscale_create(Engine& e);
void twoscale_destroy(Engine& e);
// masks: the coarse density pass's row words when the step runs the hit-mask
// pair (that pass then also set the split flags, from twoscale_split_refs),
// or null
void twoscale_enqueue(Engine& e, const uint64_t* masks);
SplitRefs twoscale_split_refs(Engine& e);
sph_status twoscale_read(Engine& e, sph_field f, void* dst, size_t bytes_out);
sph_status twoscale_counts(Engine& e, sph_counts* out);
size_t twoscale_bytes(const Engine& e);
uint8_t* twoscale_final_state(Engine& e);
// True once the sub-particle pass has run: its launches are then the same
// every step (steady state), so the two-scale step can be captured.
bool twoscale_steady(const Engine& e);
// One AdvanceSubParticles step of the current sub-particles (enqueued).
sph_status twoscale_sub_step(Engine& e);
// Sub-particle positions of the last step by sub id into out (8N float4),
// and the default / active sub id lists and counts into the view.
sph_status twoscale_export(Engine& e, float4* out, sph_render_view* view);

// Learned boundary (neural.hip).
sph_status neural_attach(Engine& e, int scene, sph_mlp* m, const float* val, size_t nval,
                         const sph_neural_config& cfg, sph_neural_report* rep);
// Frees the learned boundary; with restore, gives the engine its float
// distance copy back (it was released while every scene had a network).
sph_status neural_destroy(Engine& e, bool restore = true);
void neural_enqueue(Engine& e);
size_t neural_bytes(const Engine& e);
const float4* neural_records(const Engine& e);   // per sorted slot, or null
uint8_t* neural_need(const Engine& e);           // per sorted slot, or null
uint32_t* neural_list_count(const Engine& e);    // per scene, or null

}  // namespace sph
