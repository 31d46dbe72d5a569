/*
 * sph_oracle.h -- CPU oracle for the SPH step (TEST INFRASTRUCTURE).
 *
 * This library is the checker of the HIP implementation, never part of the
 * product: only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
 * leg load it.  It restates the reference step of CUDASPHNEW
 * (particle_simulation.cu) on the CPU, templated on the scalar type:
 *   real_kind 0 -> double (the FP64 parity oracle)
 *   real_kind 1 -> float  (FP32 transcription, used as the CPU baseline)
 * Parameter and boundary structs are the ones of include/sph.h so that the
 * same inputs can be handed to the oracle and to the HIP library.
 */
#ifndef SPH_ORACLE_H
#define SPH_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#include "../include/sph.h"

#ifdef __cplusplus
extern "C" {
#endif

/* flags reported by the key computation */
#define ORC_FLAG_OUT_OF_DOMAIN 1u
#define ORC_FLAG_NONFINITE 2u

/* oracle-only fields (double [n]) of the last acceleration pass: the
 * colour-gradient norm compared with normThresh and the wall distance
 * compared with r0 (the two switches of particle_simulation.cu:407,457) */
#define ORC_FIELD_COLOR_GRAD_NORM 100
#define ORC_FIELD_WALL_DISTANCE 101
/* colour-gradient norm of each sub-particle (sub-list order) */
#define ORC_FIELD_SUB_COLOR_GRAD_NORM 102

typedef struct orc_sim orc_sim;

/* ---- stateful simulation (storage in original-id order, as the
 * reference) ---- */
orc_sim* orc_create(const sph_params* params, const sph_boundary* boundary,
                    uint32_t n, const float* pos_xyz, const float* vel_xyz,
                    int real_kind, int threads);
/* Same, with FP64 initial state. */
orc_sim* orc_create_f64(const sph_params* params,
                        const sph_boundary* boundary, uint32_t n,
                        const double* pos_xyz, const double* vel_xyz,
                        int real_kind, int threads);
void orc_destroy(orc_sim* s);
/* n_steps of the plain step (two_scale = 0) or of AdvanceTwoScale. */
int orc_step(orc_sim* s, int n_steps, int two_scale);
/* AdvanceSubParticles (cu:2119-2143): n steps of the current sub-particles. */
int orc_step_sub(orc_sim* s, int n_steps);
/* Fields use the sph_field numbering.  Float-valued fields are written as
 * double; integer fields as int32; SPH_FIELD_STATE as uint8.  Returns the
 * number of elements written, or -1 if bytes is too small. */
long orc_read(orc_sim* s, int field, void* dst, size_t bytes);
void orc_counts(orc_sim* s, sph_counts* out);
uint32_t orc_flags(orc_sim* s);
/* Wall-clock seconds of the last orc_step call. */
double orc_last_step_seconds(orc_sim* s);

/* ---- stateless pieces (for unit tests) ---- */
/* Cell keys (compute_particle_hash, cu:89-119) of float positions, using
 * float arithmetic when real_kind = 1 (bit-exact target of the device
 * kernel) or double when real_kind = 0.  Returns the OR of the flags. */
uint32_t orc_keys(const sph_params* p, uint32_t n, const float* pos_xyz,
                  int32_t* keys, int real_kind);
/* thrust::sort_by_key of (key, id = 0..n-1) then compute_cell_start_end
 * (cu:155-193,2338-2369).  cell_start/cell_end have grid_dim product
 * entries, empty = -1. */
void orc_sort_cells(const sph_params* p, uint32_t n, const int32_t* keys,
                    int32_t* sorted_keys, int32_t* sorted_ids,
                    int32_t* cell_start, int32_t* cell_end);
/* 3D-texture lookup (normalized coordinates, linear filter, clamp) of one
 * boundary grid at world points: which = 0 distance, 1 density,
 * 2 viscosity. */
void orc_tex(const sph_boundary* b, int which, uint32_t n, const double* xyz,
             double* out);
/* O(n^2) density and acceleration (no cell grid), FP64, for checking the
 * neighbour search on small inputs. */
void orc_density_brute(const sph_params* p, const sph_boundary* b,
                       uint32_t n, const double* pos, double* rho);
void orc_accel_brute(const sph_params* p, const sph_boundary* b, uint32_t n,
                     const double* pos, const double* vel, const double* rho,
                     double* acc);

/* Sub-particle physics of one step on a given state (FP64): sub grid, the
 * density of the split children [0, nsplit_subs) and their acceleration;
 * the boundary children [nsplit_subs, nsub) take the densities rho_in.
 * Outputs: densities (nsub), accelerations (3 nsub), colour-gradient norms
 * (nsub; 1e300 where no acceleration pass ran).  Returns the key flags or
 * -1 on bad sizes. */
int orc_sub_forces(const sph_params* p, const sph_boundary* b, uint32_t nsub,
                   uint32_t nsplit_subs, const double* pos, const double* vel,
                   const double* rho_in, int threads, double* rho_out,
                   double* acc, double* cgn);

/* ---- boundary grid builder (FP64 restatement of SphInComplexShapes) ---- */
int orc_scene_boundary(const sph_scene_desc* d, double* distance,
                       double* density, double* viscosity);
void orc_scene_sdf(const sph_scene_desc* d, uint32_t n, const double* xyz,
                   double* out);

/* sizeof / offsetof of the ABI structs: params, boundary, scene,
 * scene_desc, counts, solid, then four field offsets. */
void orc_struct_layout(size_t* out10);

#ifdef __cplusplus
}
#endif

#endif
