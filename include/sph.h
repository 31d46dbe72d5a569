/*
 * sph.h -- C ABI of the MI355X (gfx950) SPH time step.
 *
 * This is the drop-in boundary for the SPH step of CUDASPHNEW
 * (CUDASPHNEW/particle_simulation.h, particle_simulation.cu).  Every entry
 * point below names the reference interface it replaces.  Plain C types
 * only: pointers, sizes and POD structs.  No exception crosses this
 * boundary; every call returns an sph_status and the handle keeps a
 * message (sph_last_error_message).
 *
 * Ownership: a handle owns all device memory it allocates.  Caller arrays
 * are copied at create time.  Pointers returned by sph_map_* are borrowed
 * and stay valid until the next step or sph_destroy on that handle.
 *
 * Threading: one HIP stream per handle.  Calls on one handle are not
 * thread-safe; distinct handles may be driven from distinct threads.  A step
 * performs no host synchronisation (so it can be captured in a hipGraph);
 * errors detected on the device (non-finite state, particle out of the cell
 * grid) are reported by the next sph_sync or read call.
 */
#ifndef SPH_H
#define SPH_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SPH_ABI_VERSION 1

/* Empty cell marker of the cell tables (particle_simulation.cu:1661). */
#define SPH_EMPTY_CELL 0xFFFFFFFFu

/* Error codes.  Replaces the C++ exceptions thrown by CUDA_SAFE_CALL /
 * THROW_EXCEPTION (util.h:21-52), which Advance swallowed
 * (particle_simulation.cu:2112-2116). */
typedef enum sph_status {
    SPH_OK = 0,
    SPH_INVALID_ARG = 1,
    SPH_OUT_OF_MEMORY = 2,
    SPH_DEVICE_ERROR = 3,
    SPH_NONFINITE_STATE = 4,
    SPH_OUT_OF_DOMAIN = 5,
    SPH_NO_DEVICE = 6,
    SPH_CAPACITY_EXCEEDED = 7
} sph_status;

/* Simulation parameters.  The first block mirrors struct SimulationParameters
 * (particle_simulation.h:54-111) field by field; the second block holds the
 * options this build adds, each with a value that reproduces the reference. */
typedef struct sph_params {
    /* --- particle_simulation.h:56-62: uniform cell grids --- */
    float grid_origin[3];
    float grid_spacing;
    int32_t grid_dim[3];
    float grid_spacing_sub;
    int32_t grid_dim_sub[3];
    /* --- particle_simulation.h:64-69 --- */
    uint32_t num_particles;
    float compact_support;          /* h */
    float compact_support_sub;      /* h_sub */
    int32_t kernel_particles;
    /* --- particle_simulation.h:72-74 --- */
    float time_step;
    float time_step_sub;
    /* --- particle_simulation.h:76-85 --- */
    float particle_mass;
    float sub_particle_mass;
    float gas_stiffness;
    float rest_density;
    float dynamic_viscosity;
    float gravity;                  /* a_y -= gravity (particle_simulation.cu:471) */
    float tension_coefficient;
    float norm_thresh;
    float fluid_volume;
    /* --- particle_simulation.h:88-98: kernel coefficients --- */
    float poly6, grad_poly6, lapl_poly6, grad_spiky, lapl_visc;
    float poly6_sub, grad_poly6_sub, lapl_poly6_sub, grad_spiky_sub,
          lapl_visc_sub;
    /* --- particle_simulation.h:102-110 --- */
    float cm_distance_thresh;
    float n_part_thresh;
    float box_dim[3];
    float box_cen[3];
    float restitution;

    /* --- options added by this build --- */
    /* 0: no collision response (reference Advance, cu:2085-2117);
     * 1: collision response of collision_handling (cu:742-834) driven by the
     *    boundary distance grid, applied inside integration. */
    int32_t collision;
    /* Two-scale split region, inclusive AABB.  The reference hard-codes
     * x in [0.2, 0.5] (cu:947): split_lo = {0.2,-inf,-inf},
     * split_hi = {0.5,+inf,+inf}. */
    float split_lo[3];
    float split_hi[3];
    /* 0: reference seeding, 8 children at radius (3m/(4 pi rho))^(1/3)
     *    along the cube diagonals (cu:1023-1090);
     * 1: children at (+-h/4, +-h/4, +-h/4) around the parent. */
    int32_t sub_placement;
    /* Capacity for active sub-particles; 0 means 8 * num_particles, the
     * reference allocation (cu:2014-2016); at most 268,435,448. */
    uint32_t max_sub_particles;
} sph_params;

/* The three boundary grids of "SPH in complex shapes"
 * (SphInComplexShapes.cpp:366-427, uploaded as 3D textures by
 * particle_simulation.cu:2592-2683,2959-3011).  Arrays are x-fastest:
 * index = i + dims[0]*(j + dims[1]*k).  The lookup follows the reference's
 * normalized-coordinate, linear-filter, clamp-address texture: a point x is
 * sampled at texel coordinate ((x-origin)/L)*dims - 0.5 with
 * L = (dims-1)*spacing. */
typedef struct sph_boundary {
    float origin[3];
    float spacing;
    int32_t dims[3];
    float rest_distance;            /* r0 (gBoundaryRestDistance, cu:46) */
    const float* distance;          /* signed, > 0 inside solids */
    const float* density;           /* wall density contribution */
    const float* viscosity;         /* wall viscosity coefficient */
} sph_boundary;

/* Initial particles plus boundary.  Replaces the host arrays filled by
 * create_particle_box (cu:2866-2917) and copied by Init (cu:1882-2005). */
typedef struct sph_scene {
    uint32_t num_particles;
    const float* positions;         /* xyz, AoS stride 3 (VD_POS_*, h:17-23) */
    const float* velocities;        /* xyz, AoS stride 3; NULL = zero */
    sph_boundary boundary;
} sph_scene;

typedef struct sph_handle sph_handle;

/* Fields readable with sph_read_field.  Per-particle fields are returned in
 * original-id order (the reference never permutes its storage, so its
 * arrays are in that order).  Sorted fields are in cell-sorted order. */
typedef enum sph_field {
    SPH_FIELD_POSITION = 0,     /* float xyz [3n]   (VD_POS_*)            */
    SPH_FIELD_VELOCITY = 1,     /* float xyz [3n]   (SD_VEL0_*)           */
    SPH_FIELD_DENSITY = 2,      /* float [n]        (SD_DENSITY)          */
    SPH_FIELD_PRESSURE = 3,     /* float [n]        (SD_PRESSURE)         */
    SPH_FIELD_ACCEL = 4,        /* float xyz [3n]   (SD_ACC_*)            */
    SPH_FIELD_STATE = 5,        /* uint8 [n]        (mParticleStates)     */
    SPH_FIELD_SORTED_KEYS = 6,  /* int32 [n]   (mParticleHashListDevPtr)  */
    SPH_FIELD_SORTED_IDS = 7,   /* int32 [n]   (mParticleIdsDevPtr)       */
    SPH_FIELD_CELL_START = 8,   /* int32 [cells]  (mCellStartListDevPtr)  */
    SPH_FIELD_CELL_END = 9,     /* int32 [cells]  (mCellEndListDevPtr)    */
    SPH_FIELD_SUB_POSITION = 10,/* float xyz [3*nsub], sub-list order     */
    SPH_FIELD_SUB_VELOCITY = 11,/* float xyz [3*nsub]                     */
    SPH_FIELD_SUB_DENSITY = 12, /* float [nsub]                           */
    SPH_FIELD_SUB_PRESSURE = 13,/* float [nsub]                           */
    SPH_FIELD_SUB_ACCEL = 14,   /* float xyz [3*nsub]                     */
    SPH_FIELD_SUB_IDS = 15,     /* int32 [nsub]: 8*parent_id + child      */
    SPH_FIELD_SPLIT_IDS = 16,   /* int32 [nsplit]  (mParticleIdsSplit)    */
    SPH_FIELD_BOUNDARY_IDS = 17,/* int32 [nboundary]                      */
    SPH_FIELD_DEFAULT_IDS = 18, /* int32 [ndefault] (mParticleIdsDefault) */
    SPH_FIELD_SUB_SORTED_KEYS = 19, /* int32 [nsub] (mSubParticleHashs)   */
    SPH_FIELD_SUB_SORTED_IDS = 20,  /* int32 [nsub] sub ids, sorted       */
    SPH_FIELD_WALL_SOURCE = 21, /* uint8 [n]: boundary source of the last
                                   step (0 none, 1 network, 2 grid) */
    /* Sub-particle cell tables (mSubParticleCellStartIdsDevPtr /
     * mSubParticleCellEndIdsDevPtr, particle_simulation.h:269, built by
     * cu:2371-2389): int32 [sub cells], first / one past the last slot of
     * each sub cell in SUB_SORTED_* order, -1 = empty. */
    SPH_FIELD_SUB_CELL_START = 22,
    SPH_FIELD_SUB_CELL_END = 23,
    /* float xyz [3*nsub], sub-list order: the sub-particle state the last
     * sub pass started from -- the positions its sub-grid hash keyed and
     * the velocities its forces used (the seeded state after a two-scale
     * step, initialize_sub_particles cu:1023-1090; the state at the start
     * of the step after a sub-particle step). */
    SPH_FIELD_SUB_INPUT_POSITION = 24,
    SPH_FIELD_SUB_INPUT_VELOCITY = 25
} sph_field;

/* Counts of the two-scale classes (GetNumParticles, GetNumParticlesDefault,
 * GetNumSubParticles, GetNumSubParticles{Regular,Boundary},
 * particle_simulation.cu:2227-2316). */
typedef struct sph_counts {
    uint32_t num_particles;
    uint32_t num_split;
    uint32_t num_boundary;
    uint32_t num_default;
    uint32_t num_sub;
    uint32_t num_steps;             /* GetNumTimesSteps (cu:2245) */
} sph_counts;

/* ---- lifecycle ---------------------------------------------------------
 * sph_create replaces the factory + Init + Bind sequence
 * (particle_simulation.h:174, cu:1745, cu:1882, cu:2078).  Parameters are
 * validated; the handle owns copies of everything.  At most 268,435,455
 * particles per handle (or batch): the neighbour gathers address particles
 * with 32-bit byte offsets. */
sph_status sph_create(const sph_params* params, const sph_scene* scene,
                      sph_handle** out);
/* ~ParticleSimulation (cu:1689-1735). */
sph_status sph_destroy(sph_handle* h);

/* ---- stepping ----------------------------------------------------------
 * sph_step: n_steps of Advance (cu:2085-2117) restricted to the coarse
 * physics (hash, sort, cell ranges, density/pressure, acceleration,
 * integration [+ collision]).  Asynchronous on the handle's stream. */
sph_status sph_step(sph_handle* h, int32_t n_steps);
/* sph_step_two_scale: n_steps of AdvanceTwoScale (cu:2145-2182): the coarse
 * step plus state classification, device-side compaction, sub-particle
 * seeding, projection, sub-grid sort, sub density/acceleration, sub
 * integration and collision.  No host synchronisation inside. */
sph_status sph_step_two_scale(sph_handle* h, int32_t n_steps);
/* sph_step_sub_particles: n_steps of AdvanceSubParticles (cu:2119-2143):
 * the sub-particles of the last two-scale step advance alone (sub grid,
 * sub density, projection for the boundary subs, sub acceleration,
 * integration and collision; no reseeding, the coarse particles fixed).
 * Projection pairs the coarse positions as the last step left them with
 * that step's cell ranges and densities, as the reference does.
 * SPH_INVALID_ARG before the first two-scale step. */
sph_status sph_step_sub_particles(sph_handle* h, int32_t n_steps);
/* Waits for the handle's stream; reports device-side errors. */
sph_status sph_sync(sph_handle* h);

/* ---- parameters --------------------------------------------------------
 * Replaces SetNPartThresh / Increase/DecreaseCmDistanceThresh + Bind
 * (cu:2255-2276).  Takes effect at the next step.  Grid geometry and the
 * particle count may not change. */
sph_status sph_set_params(sph_handle* h, const sph_params* params);
sph_status sph_get_params(const sph_handle* h, sph_params* out);

/* ---- read back / map ---------------------------------------------------
 * sph_read_positions replaces the GL-mapped VBO read
 * (GetGLParticleVertexBufferObject + map, cu:2278,2685): xyz in
 * original-id order. */
sph_status sph_read_positions(sph_handle* h, float* xyz, size_t n);
/* Generic field read (GetParticleState cu:2238, SaveInfoTable cu:2709,
 * SaveParticleInfo cu:2785).  bytes must match the field size exactly. */
sph_status sph_read_field(sph_handle* h, sph_field field, void* dst,
                          size_t bytes);
/* Zero-copy access: device pointers to the cell-sorted float4 positions
 * (x, y, z, density) and the original id of each sorted slot.  Valid until
 * the next step or destroy. */
sph_status sph_map_positions(sph_handle* h, const float** dev_pos4,
                             const uint32_t** dev_orig_id);
/* Render interop (ParticleSimulation::map / unmap, cu:2685-2707, over the
 * vertex and index buffers of cu:1926-1932 and 2009-2035): device pointers
 *   positions      float4 (x, y, z, density) per particle, by original id
 *                  [num_particles]
 *   sub_positions  float4 (x, y, z, density) per sub-particle, by
 *                  sub-particle id 8 * parent + child [8 * num_particles]
 *   default_ids    original ids of the particles in default state
 *                  [num_default]
 *   sub_ids        ids of the active sub-particles, regular (children of
 *                  split particles) first, then boundary
 *                  [num_sub_regular + num_sub_boundary]
 * A renderer draws default_ids out of positions and sub_ids out of
 * sub_positions (as twoscalestatesub.cpp:96-113 does with GL element
 * buffers); before the first two-scale step the last three are NULL / 0 and
 * all particles are drawn from positions.  The reference's vertex buffer has
 * 3 floats per particle; here 4, the fourth being the density.
 * sph_map_render fills the buffers on the handle's stream and waits for
 * them; the pointers stay valid until the next step or destroy. */
typedef struct sph_render_view {
    const float* positions;
    const float* sub_positions;
    const uint32_t* default_ids;
    const uint32_t* sub_ids;
    uint32_t num_particles;
    uint32_t num_default;
    uint32_t num_sub_regular;
    uint32_t num_sub_boundary;
} sph_render_view;
sph_status sph_map_render(sph_handle* h, sph_render_view* out);
/* unmap: ends a map.  The library keeps no mapping state (its buffers are
 * its own, not a graphics API's), so this only checks the handle. */
sph_status sph_unmap_render(sph_handle* h);
sph_status sph_get_counts(sph_handle* h, sph_counts* out);
/* Debug text dumps in the reference's column formats (iostream setw):
 * SaveInfoTable (cu:2709-2780): "Number of particles N", header, then one
 * row per cell index i: i, and for i < N the sorted id and hash of slot i,
 * then the cell's start and end (blank when empty).  SaveParticleInfo
 * (cu:2782-2852): header and one row per particle id: position, density,
 * pressure, acceleration, 20 characters per column, default float format.
 * Scene 0 of the handle; host buffers, a device-to-host copy of the fields. */
sph_status sph_save_info_table(sph_handle* h, const char* filename);
/* extractSurfaceParticles: declared by the reference (particle_simulation.h:224)
 * but never defined; its inputs are -- the thresholds cmDistanceThresh and
 * nPartTresh (h:103-104, set to 0.5 and 20 by Example01, cu:1811-1812) and
 * the neighbourhood centre-of-mass sums of compute_particle_acceleration_ifsurf
 * (cu:370-380, 1550-1553: the particle itself plus every neighbour with
 * 0 < r <= h).  Definition used here, on the positions the last step's
 * forces were computed from (the state the reference's kernel would see):
 *   n_i = 1 + #{j : 0 < |x_j - x_i| <= h},
 *   d_i = | sum_j (x_j - x_i) | / n_i   (centre of mass minus x_i),
 *   surface  <=>  n_i < nPartTresh  or  d_i > cmDistanceThresh * h.
 * is_surface: one byte per particle, by original id (0/1); num_surface:
 * count of flagged particles (may be NULL).  Needs one step first. */
sph_status sph_extract_surface(sph_handle* h, uint8_t* is_surface, uint32_t* num_surface);
sph_status sph_save_particle_info(sph_handle* h, const char* filename);
size_t sph_num_cells(const sph_handle* h);
/* GetSizeMemoryGPU (particle_simulation.h:169; not defined in the
 * reference): device bytes owned by the handle. */
size_t sph_device_bytes(const sph_handle* h);
/* Message of the last error on this handle ("" if none); NULL handle gives
 * the last error of sph_create or sph_step_noalloc on this thread. */
const char* sph_last_error_message(const sph_handle* h);

/* ---- timing ------------------------------------------------------------
 * Per-stage device time, averaged per step since sph_enable_timing, measured
 * with HIP events recorded on the handle's stream around every stage
 * (CudaTimer, util.h:193-213, re-expressed); the events are resolved when
 * sph_stage_times is called, so timing adds no host synchronisation to a
 * step.  Stage names: see sph_stage_name.  Sub-particle steps
 * (sph_step_sub_particles) add their stage times to the same totals but are
 * not counted as steps: the averages are per coarse step. */
#define SPH_MAX_STAGES 16
sph_status sph_enable_timing(sph_handle* h, int32_t enable);
sph_status sph_stage_times(sph_handle* h, float* ms, int32_t* count);
const char* sph_stage_name(int32_t stage);
/* Device time of n steps measured with HIP events recorded on the handle's
 * stream around the steps (no host work inside the timed region). */
/* Replays the steady-state plain step from a captured hipGraph (captured on
 * the first eligible step, re-captured after any change of parameters,
 * kernel variants or learned boundary); steps with per-stage timing on and
 * two-scale steps run eagerly.  No reference counterpart (the reference
 * launches every kernel of every step); as sph_batch_use_graph. */
sph_status sph_use_graph(sph_handle* h, int32_t enable);
sph_status sph_time_steps(sph_handle* h, int32_t n_steps, int32_t two_scale,
                          float* ms);
/* ---- kernel tuning -----------------------------------------------------
 * Variants of the density and force passes; results agree to FP32 summation
 * order.  density_lanes / force_lanes: lanes of a wavefront that share one
 * particle in the neighbour loops (1 = one thread per particle; 2, 4, 8, 16 =
 * cooperative variants).  hit_masks = 1: the density pass records the force
 * neighbours of each particle as per-row bit masks (180 bytes per particle
 * of extra device memory: 72 for rows of up to 32 candidates, the rest for
 * the longer rows of a compressed flow) and the force pass runs the pair
 * terms from them through a per-thread hit list of hit_list entries (32, 40,
 * 48 or 64) in LDS, filled as often as a particle's neighbours need;
 * the lanes fields are then unused.  No reference counterpart: the reference
 * fixes one thread per particle and a block size of 256
 * (particle_simulation.cu:1385-1556, 2943-2957); these generalise its launch
 * configuration. */
typedef struct sph_tuning {
    int32_t density_lanes;
    int32_t force_lanes;
    int32_t hit_masks;
    int32_t hit_list;
    /* threads per workgroup of the one-lane and hit-mask variants of both
     * passes: 128 or 256 (0 = 256); the cooperative variants keep 256 */
    int32_t block;
} sph_tuning;
sph_status sph_set_tuning(sph_handle* h, const sph_tuning* t);
sph_status sph_get_tuning(const sph_handle* h, sph_tuning* out);
/* Times every variant on the current state (needs one step first; each
 * kernel reps times), keeps the fastest, re-runs the last step's density and
 * force with it; out may be NULL. */
sph_status sph_autotune(sph_handle* h, int32_t reps, sph_tuning* out);
/* Re-runs sph_autotune (reps 2) before every steps-th step of sph_step and
 * sph_step_two_scale, on the live state; 0 (the default) turns it off.  The
 * best variants change as a flow develops: a compressing dam break picks the
 * hit-mask pair near rest density and the one-lane pair once its particles
 * have two to three times the neighbours.  Each re-tune times every variant
 * on the live state (18-56 ms at 1M particles, more as the flow compresses),
 * synchronises the stream and, with graphs on, re-captures the step: use an
 * interval of thousands of steps.  No reference counterpart (the reference
 * has one launch configuration). */
sph_status sph_set_autotune_interval(sph_handle* h, int32_t steps);

/* ---- batched scenes ----------------------------------------------------
 * Many independent scenes advanced together: one launch per stage over all
 * scenes, scenes kept in disjoint key ranges of one segmented sort.  The
 * per-scene interface is that of sph_create.  No reference counterpart (the
 * reference binds one simulation at a time, particle_simulation.h:122-125);
 * this generalises sph_create/sph_step to a list of scenes. */
typedef struct sph_batch sph_batch;
sph_status sph_batch_create(int32_t n_scenes, const sph_params* params,
                            const sph_scene* scenes, sph_batch** out);
sph_status sph_batch_destroy(sph_batch* b);
sph_status sph_batch_step(sph_batch* b, int32_t n_steps);
sph_status sph_batch_sync(sph_batch* b);
/* Positions of scene s in that scene's original-id order. */
sph_status sph_batch_read_field(sph_batch* b, int32_t scene, sph_field field,
                                void* dst, size_t bytes);
sph_status sph_batch_time_steps(sph_batch* b, int32_t n_steps, float* ms);
/* Per-stage device times of a batch, as sph_enable_timing/sph_stage_times. */
sph_status sph_batch_enable_timing(sph_batch* b, int32_t enable);
sph_status sph_batch_stage_times(sph_batch* b, float* ms, int32_t* count);
/* hipGraph capture of one batched step (0 = launch kernels directly). */
sph_status sph_batch_use_graph(sph_batch* b, int32_t enable);
/* Kernel tuning of a batch, as sph_set_tuning/sph_get_tuning/sph_autotune. */
sph_status sph_batch_set_tuning(sph_batch* b, const sph_tuning* t);
sph_status sph_batch_get_tuning(const sph_batch* b, sph_tuning* out);
sph_status sph_batch_autotune(sph_batch* b, int32_t reps, sph_tuning* out);
sph_status sph_batch_set_autotune_interval(sph_batch* b, int32_t steps);
uint64_t sph_batch_num_particles(const sph_batch* b);
const char* sph_batch_last_error_message(const sph_batch* b);

/* ---- scene construction (host only, no device needed) ------------------
 * Replaces Example01 (cu:1745-1817), create_particle_box (cu:2866-2917),
 * set_simulation_domain (cu:2919-2941) and the boundary grid builder
 * SphInComplexShapes (SphInComplexShapes.cpp:34-64,203-313,366-427). */
#define SPH_MAX_OBSTACLES 8
typedef enum sph_solid_kind {
    SPH_SOLID_NONE = 0,
    SPH_SOLID_SPHERE = 1,   /* center, radius = a[0]                     */
    SPH_SOLID_CONE = 2,     /* base center, base radius a[0], height a[1],
                               apex at center + (0, a[1], 0)             */
    SPH_SOLID_BOX = 3       /* center, half extents a[0..2]              */
} sph_solid_kind;

typedef struct sph_solid {
    int32_t kind;
    float center[3];
    float a[3];
} sph_solid;

typedef struct sph_scene_desc {
    /* fluid block: lattice[0]*lattice[1]*lattice[2] particles at
     * origin + spacing*(i,j,k), jittered by U(-jitter, jitter) per axis
     * from a counter-based hash of (seed, particle, axis). */
    uint32_t lattice[3];
    float lattice_origin[3];
    float lattice_spacing;
    float jitter;
    uint64_t seed;
    /* container box (fluid inside), obstacles (solid) */
    float box_lo[3];
    float box_hi[3];
    int32_t num_obstacles;
    sph_solid obstacles[SPH_MAX_OBSTACLES];
    /* physics */
    float h;
    float rest_density;
    float particle_mass;
    float gas_stiffness;
    float viscosity;
    float gravity;
    float tension;
    float norm_thresh;
    float dt;
    float dt_sub;
    /* cell grid: set_simulation_domain(cell_lo, cell_hi, h, h/2)
     * (cu:2919-2941): dims = int((hi-lo)/spacing + 0.5) */
    float cell_lo[3];
    float cell_hi[3];
    /* boundary grid */
    float boundary_origin[3];
    float boundary_spacing;
    int32_t boundary_dims[3];
    float rest_distance;            /* r0 */
    float boundary_support;         /* cs of the grid builder */
    float ghost_spacing;            /* 0 = (m/rho0)^(1/3) (cu:2598) */
    /* options copied into sph_params */
    int32_t collision;
    float restitution;
    float split_lo[3];
    float split_hi[3];
    int32_t sub_placement;
} sph_scene_desc;

/* Named workloads. */
typedef enum sph_config_id {
    SPH_CONFIG_EXAMPLE01 = 0,   /* particle_simulation.cu:1745-1817 */
    SPH_CONFIG_DAM16K = 1,      /* 16,384 particles, 0.64^3 box, 64^3 grid */
    SPH_CONFIG_DAM1M = 2,       /* 1,048,576 particles, sphere + cone */
    SPH_CONFIG_TWO_SCALE = 3,   /* 262,144 coarse + sub-particles */
    SPH_CONFIG_BATCH = 5        /* one scene of the 64-scene batch */
} sph_config_id;

/* Fills desc for a named workload.  For SPH_CONFIG_BATCH, index selects the
 * scene (0..63) of the batch drawn from seed. */
sph_status sph_scene_config(int32_t config, int32_t index, uint64_t seed,
                            sph_scene_desc* desc);
/* Derives sph_params (kernel coefficients, grids) from a description. */
sph_status sph_scene_params(const sph_scene_desc* desc, sph_params* out);
/* Generates particle positions (xyz AoS).  Returns the count in *n; pass
 * pos = NULL to query the count. */
sph_status sph_scene_particles(const sph_scene_desc* desc, float* pos,
                               uint32_t capacity, uint32_t* n);
/* Builds the three boundary grids (dims from desc).  Follows
 * SphInComplexShapes::SetBox for the container box and extends the same
 * ghost-particle construction to the obstacle solids.  Each output holds
 * boundary_dims[0]*[1]*[2] floats. */
sph_status sph_scene_boundary(const sph_scene_desc* desc, float* distance,
                              float* density, float* viscosity);
/* Signed distance of the union of solids at n points (FP32 analytic),
 * > 0 inside a solid. */
sph_status sph_scene_sdf(const sph_scene_desc* desc, const float* xyz,
                         size_t n, float* out);

/* An obstacle given by a sampled signed distance, e.g. a mesh's boundary
 * map (sph_boundary_map_generate_on).  No reference counterpart: the
 * reference generates mesh maps (main.cpp:91-105) but its grid builder
 * only knows boxes.  Nodes at lo + dx (i, j, k), x fastest; distance
 * negative inside.  Inside a cell whose eight nodes hold distances the
 * solid's distance is trilinear; elsewhere a point is taken to be at least
 * `band` from the surface, on the side of its nearest node (nodes without
 * a distance take the side of the band region they border), and points off
 * the grid are outside.  Usable when band >= support + ghost spacing +
 * sqrt(3) dx and the grid covers [solid_lo, solid_hi] grown by support +
 * ghost spacing (else SPH_INVALID_ARG). */
typedef struct sph_map_solid {
    int32_t dims[3];
    float lo[3];
    float dx;
    float band;                 /* distances are exact within band (maxDist) */
    const float* distance;      /* dims[0]*dims[1]*dims[2] */
    const uint8_t* present;     /* 1 = node holds a distance; NULL = all do */
    float solid_lo[3];          /* the solid's bounding box: the ghost      */
    float solid_hi[3];          /* lattice starts at solid_lo - (cs + gsp)  */
} sph_map_solid;
/* sph_scene_boundary with map solids added to the desc's obstacles (their
 * ghosts after the analytic obstacles').  num_maps = 0 gives
 * sph_scene_boundary's grids bit for bit. */
sph_status sph_scene_boundary_maps(const sph_scene_desc* desc,
                                   const sph_map_solid* maps, int32_t num_maps,
                                   float* distance, float* density,
                                   float* viscosity);

/* Example01's own grid builder call (cu:2592-2612): box of half extents
 * (0.7,0.5,0.3) on the grid (-0.8,-0.6,-0.4)..(0.8,0.6,0.4), spacing h/4.
 * Writes dims and, if the pointers are non-NULL, the node data. */
sph_status sph_example01_boundary(int32_t dims[3], float origin[3],
                                  float* spacing, float* rest_distance,
                                  float* distance, float* density,
                                  float* viscosity);

/* Training samples of the boundary's signed distance: n points, a fraction
 * near_frac of them within near_width of the surface (|sdf| < near_width,
 * by rejection), the rest uniform over the boundary grid's box; each sample
 * is (x, y, z, sdf) with sdf as sph_scene_sdf.  Seeded, deterministic. */
sph_status sph_scene_sdf_samples(const sph_scene_desc* desc, uint64_t seed,
                                 size_t n, float near_frac, float near_width,
                                 float* xyzd);

/* The same sample points, each labelled with the boundary grid's distance
 * texture at the point: the field the step samples (node distances within
 * the builder's support, clamped to +-r0, -r0 elsewhere; lookup convention
 * of particle_simulation.cu:243-245, 3003-3007). */
sph_status sph_scene_grid_samples(const sph_scene_desc* desc, uint64_t seed,
                                  size_t n, float near_frac, float near_width,
                                  float* xyzd);

/* sph_scene_sdf with the map solids in the union. */
sph_status sph_scene_sdf_maps(const sph_scene_desc* desc,
                              const sph_map_solid* maps, int32_t num_maps,
                              const float* xyz, size_t n, float* out);
/* sph_scene_sdf_samples / sph_scene_grid_samples with the map solids in
 * the union (the training and validation samples of a learned boundary for
 * a scene whose obstacles come from meshes). */
sph_status sph_scene_sdf_samples_maps(const sph_scene_desc* desc,
                                      const sph_map_solid* maps,
                                      int32_t num_maps, uint64_t seed,
                                      size_t n, float near_frac,
                                      float near_width, float* xyzd);
sph_status sph_scene_grid_samples_maps(const sph_scene_desc* desc,
                                       const sph_map_solid* maps,
                                       int32_t num_maps, uint64_t seed,
                                       size_t n, float near_frac,
                                       float near_width, float* xyzd);

/* ---- learned boundary SDF ------------------------------------------------
 * No reference counterpart: a compact stand-in for the distance texture
 * (particle_simulation.cu:48-53, built by SphInComplexShapes.cpp:203-313).
 * MLP over Fourier features of the normalized position
 * u = 2 (x - lo)/(hi - lo) - 1: features [u, sin(pi f_k u_a),
 * cos(pi f_k u_a)] for k < 16, f_k = 2^(7k/15) (99, zero-padded to 128),
 * then 128 -> 64 -> 64 -> 64 (SiLU) -> 1 with output y.  The distance is
 * d = band * atanh(y) (band > 0: the network learns tanh(d / band), which
 * keeps the resolution where the boundary terms act, |d| of a few h, and
 * saturates far from the surface) or d = out_scale * y (band = 0).
 * Products run on BF16 MFMA with FP32 accumulation; the master parameters,
 * biases, output layer and Adam state are FP32.
 * Parameter vector (SPH_MLP_NUM_PARAMS floats): W1[128][64] (row = feature),
 * b1[64], W2[64][64], b2[64], W3[64][64], b3[64], w4[64], b4. */
#define SPH_MLP_NUM_PARAMS 16641
typedef struct sph_mlp sph_mlp;
typedef struct sph_mlp_config {
    float box_lo[3], box_hi[3];   /* normalization box */
    float out_scale;              /* metres per unit of output (band = 0) */
    float band;                   /* tanh compression width in metres, or 0 */
    uint64_t seed;                /* weight initialisation */
    float lr, beta1, beta2, eps;  /* Adam */
    float lr_min;                 /* cosine decay of lr to lr_min over */
    int32_t decay_steps;          /* decay_steps steps (0 = constant lr) */
} sph_mlp_config;
/* box, out_scale = half the largest extent, band 0, seed 1, Adam (1e-3,
 * 0.9, 0.999, 1e-8), constant lr. */
sph_status sph_mlp_default_config(const float box_lo[3], const float box_hi[3],
                                  sph_mlp_config* out);
/* Xavier-uniform weights from cfg->seed, zero biases. */
sph_status sph_mlp_create(const sph_mlp_config* cfg, sph_mlp** out);
sph_status sph_mlp_destroy(sph_mlp* m);
sph_status sph_mlp_get_config(const sph_mlp* m, sph_mlp_config* out);
/* FP32 parameters (resets the Adam state and step count on set). */
sph_status sph_mlp_set_params(sph_mlp* m, const float* params, size_t n);
sph_status sph_mlp_get_params(sph_mlp* m, float* params, size_t n);
/* Training set of n (x, y, z, sdf) samples, copied to the device; batches
 * are taken in order, wrapping around. */
sph_status sph_mlp_set_samples(sph_mlp* m, const float* xyzd, size_t n);
/* Gradient of the loss (1/B) sum (y - t(sdf))^2, t(d) = tanh(d / band)
 * (or d / out_scale), over the batch of
 * samples [first, first + batch) (wrapping), without updating; batch is a
 * multiple of 64. */
sph_status sph_mlp_gradient(sph_mlp* m, uint32_t first, int32_t batch,
                            float* grad, float* loss);
/* steps Adam steps on consecutive batches; losses (optional) receives the
 * loss of every step. */
sph_status sph_mlp_train(sph_mlp* m, int32_t steps, int32_t batch,
                         float* losses);
/* Device time of `steps` training steps (HIP events). */
sph_status sph_mlp_time_train(sph_mlp* m, int32_t steps, int32_t batch,
                              float* ms);
/* Distance and its gradient at n points: dgrad[4i..4i+3] = (d, dd/dx,
 * dd/dy, dd/dz). */
sph_status sph_mlp_eval(sph_mlp* m, const float* xyz, size_t n, float* dgrad);
const char* sph_mlp_last_error_message(const sph_mlp* m);

/* ---- learned boundary in the step -----------------------------------------
 * The network stands in for the distance grid: every cell of the scene's
 * cell grid takes its wall terms from
 *   none    -- every texel the cell's particles can reach holds the grid's
 *              far-field values (distance -r0, no wall density or
 *              viscosity): no wall term, no lookup (exactly what the grid
 *              gives there);
 *   network -- per step, the central difference of the network's
 *              distance at +-spacing along each axis (the stencil of the
 *              grid's normal, cu:429-439, applied to the network's field:
 *              six forward passes on MFMA) gives the normal of the wall
 *              penalty and of the collision response in place of the grid's
 *              central differences (12 trilinear lookups); the network is
 *              trained on the distance texture (sph_scene_grid_samples).  The distance itself, which
 *              switches the wall term (distWall < r0, cu:422) and sets a
 *              penalty of stiffness rho/dt^2, and the wall density and
 *              viscosity (which depend on each wall's ghost lattice, not on
 *              the distance alone) stay on the grid;
 *   grid    -- the grid lookup, where validation failed: bricks of `brick`^3
 *              cells whose validation samples (labelled as in
 *              sph_scene_grid_samples) within `band` of the surface have an
 *              RMS error > tolerance.
 * The network must outlive the attachment (it is read every step).  No
 * reference counterpart. */
typedef struct sph_neural_config {
    float tolerance;   /* validation bar, metres (e.g. 0.05 h) */
    float band;        /* |sdf| below which the bar applies (e.g. 2 h) */
    /* bar on the wall normal (0 = not checked): RMS over the validation
     * samples of a brick in the wall-penalty zone (label D > -r0) of
     * w |G_network - G_grid|, the two central differences at +-spacing in
     * world units (the network's normal and the grid's that it replaces;
     * an exact signed distance has |G| = 1), w = min((r0 + D)/r0, 1) the
     * penetration that scales the penalty force (cu:422-445): the error of
     * the penalty force relative to its value at a penetration of r0 */
    float normal_tolerance;
    int32_t brick;     /* validation brick edge, cells */
} sph_neural_config;
typedef struct sph_neural_report {
    uint64_t cells_none, cells_neural, cells_grid;
    uint32_t bricks_checked, bricks_fallback;
    float max_error, mean_error;   /* over validation samples within band */
    /* a brick falls back when the RMS error of its samples within band
     * exceeds tolerance, or the RMS normal error exceeds normal_tolerance */
    uint64_t samples_in_band;
    float mean_normal_error;       /* mean w |G_network - G_grid| */
    uint32_t bricks_fallback_normal;   /* of bricks_fallback: normal only */
} sph_neural_report;
/* val_xyzd: n_val validation samples (x, y, z, sdf), e.g. from
 * sph_scene_sdf_samples with another seed than the training set. */
sph_status sph_attach_neural_boundary(sph_handle* h, sph_mlp* net,
                                      const float* val_xyzd, size_t n_val,
                                      const sph_neural_config* cfg,
                                      sph_neural_report* report);
sph_status sph_detach_neural_boundary(sph_handle* h);
/* Learned boundary for scene `scene` of a batch, as
 * sph_attach_neural_boundary. */
sph_status sph_batch_attach_neural_boundary(sph_batch* b, int32_t scene,
                                            sph_mlp* net, const float* val_xyzd,
                                            size_t n_val,
                                            const sph_neural_config* cfg,
                                            sph_neural_report* report);

/* ---- mesh -> signed distance field ------------------------------------------
 * BoundaryMap (boundary_map.h:37-122, boundary_map.cpp): the signed
 * distance of a triangle mesh on a grid over the mesh's bounding box, with
 * the unit normal towards the node, stored where the distance is within
 * maxDist = max(restDistance, compactSupport) of the surface; negative
 * inside.  Generated on the GPU. */
typedef struct sph_mesh {
    const float* vertices;      /* xyz per vertex */
    uint32_t num_vertices;
    const uint32_t* faces;      /* three 0-based vertex indices per triangle */
    uint32_t num_faces;
} sph_mesh;
/* BoundaryMapConfiguration (boundary_map.h:25-31). */
typedef struct sph_boundary_map_config {
    float rest_distance, compact_support, dx;
} sph_boundary_map_config;
/* OBJ reader (TriangleMesh(filename), triangle_mesh.cpp): "v x y z" and
 * "f a b c" lines (1-based; polygons fanned).  Two calls: with NULL arrays
 * (or too small capacities in *num_vertices / *num_faces) it returns the
 * counts. */
sph_status sph_obj_read(const char* path, float* vertices,
                        uint32_t* num_vertices, uint32_t* faces,
                        uint32_t* num_faces);
/* TriangleMesh::fit: scale per axis about the bounding-box centre to the
 * target box, then move the box corner to lo. */
sph_status sph_mesh_fit(float* vertices, uint32_t num_vertices,
                        const float lo[3], const float hi[3]);
/* AddCanvas domain (boundary_map.cpp:60-86): dims = ceil((v2 - v1)/dx) + 1,
 * lo = v1, hi = v1 + (dims - 1) dx. */
sph_status sph_boundary_map_domain(const sph_mesh* mesh, float dx,
                                   int32_t dims[3], float lo[3], float hi[3]);
/* AddCanvas + computeSignedDistanceField (boundary_map.cpp:369-490) on the
 * GPU: dense outputs over the domain (distance; normal xyz; present = 1
 * where a distance <= maxDist was found; absent nodes hold
 * (restDistance, 0, 0, 0) as after BoundaryMap::Load).  ms (optional):
 * device time of the two kernels. */
sph_status sph_boundary_map_generate(const sph_mesh* mesh,
                                     const sph_boundary_map_config* cfg,
                                     float* distance, float* normal,
                                     uint8_t* present, float* ms);
/* The same field on a caller-given grid: dims[0]*dims[1]*dims[2] nodes at
 * lo + dx (i, j, k), dx = cfg->dx.  No reference counterpart (AddCanvas
 * always takes the mesh's bounding box); a scene obstacle needs the grid
 * to extend past the box so that the band is whole on every side
 * (sph_map_solid).  Triangles whose grown bounds miss the grid are
 * skipped. */
sph_status sph_boundary_map_generate_on(const sph_mesh* mesh,
                                        const sph_boundary_map_config* cfg,
                                        const int32_t dims[3],
                                        const float lo[3], float* distance,
                                        float* normal, uint8_t* present,
                                        float* ms);
/* BoundaryMap::Save / Load text format (boundary_map.cpp:161-279). */
sph_status sph_boundary_map_save(const char* path,
                                 const sph_boundary_map_config* cfg,
                                 const int32_t dims[3], const float lo[3],
                                 const float hi[3], const float* distance,
                                 const float* normal, const uint8_t* present);
sph_status sph_boundary_map_load(const char* path, sph_boundary_map_config* cfg,
                                 int32_t dims[3], float lo[3], float hi[3],
                                 float* distance, float* normal,
                                 uint8_t* present);

/* ---- allocation-free step ---------------------------------------------------
 * No reference counterpart.  For callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc): n_steps of sph_step's plain step (hash, counting sort by cell,
 * reorder, density/pressure, acceleration + integration [+ collision]) with
 * sph_step's default kernel variants, on a particle set held in caller
 * device memory.  The call allocates nothing, copies nothing between host
 * and device and does not synchronise: kernels only, on `stream` (a
 * hipStream_t; NULL = the null stream).  Every other buffer (keys, orders,
 * sorted copies, cell table, sort counts: SPH_NOALLOC_SCRATCH_BYTES) lives in
 * one static device array of the library, one per GPU, held until the
 * process ends.  All calls on one GPU share it: each call clears what must
 * start cleared, but calls that overlap on one GPU (two streams, two host
 * threads) race and give undefined results.  The first call on a device
 * looks the array's address up (not a stream operation): make it outside
 * graph capture.
 *   positions   device float4 [n] (x, y, z, w), 16-byte aligned; in/out
 *   velocities  device float4 [n] (vx, vy, vz, w); in/out
 *               (particle i stays at index i; after the call w of both is
 *               the density of the last step)
 *   boundary    the boundary grids of sph_boundary, in device memory:
 *               grid = float4 per node (distance, wall density, wall
 *               viscosity, unused), distance = float per node, x fastest
 *   status_flags  optional device uint32: set to 0, then OR-ed with 1
 *               (particle outside the cell grid) and 2 (non-finite state)
 * Limits: n <= SPH_NOALLOC_MAX_PARTICLES, cells <= SPH_NOALLOC_MAX_CELLS,
 * else SPH_INVALID_ARG (message via sph_last_error_message(NULL)).  The
 * result equals sph_create + sph_step(n_steps) + reading the fields by
 * original id, bit for bit. */
#define SPH_NOALLOC_MAX_PARTICLES 262144u
#define SPH_NOALLOC_MAX_CELLS 1048576u
#define SPH_NOALLOC_SCRATCH_BYTES 39848960u
typedef struct sph_device_boundary {
    float origin[3];
    float spacing;
    int32_t dims[3];
    float rest_distance;
    const float* grid;              /* float4 per node */
    const float* distance;          /* float per node */
} sph_device_boundary;
sph_status sph_step_noalloc(const sph_params* params,
                            const sph_device_boundary* boundary,
                            uint32_t num_particles, float* positions,
                            float* velocities, int32_t n_steps,
                            uint32_t* status_flags, void* stream);

/* ---- library info ----------------------------------------------------- */
int32_t sph_abi_version(void);
/* Number of visible HIP devices (0 on a machine without a GPU). */
int32_t sph_device_count(void);

#ifdef __cplusplus
}
#endif

#endif /* SPH_H */
