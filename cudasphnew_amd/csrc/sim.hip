// sim.hip -- host orchestration and the C ABI of the SPH step.
//
// One Engine advances one or many scenes.  The reference's stage sequence
// (ParticleSimulation::Advance, particle_simulation.cu:2085-2117) becomes:
//   [hash]  clear  sort  cells  reorder  density  force(+integrate)
// all enqueued on the handle's stream without host synchronisation; counts
// stay on the device.  Errors cross the ABI as sph_status codes.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iomanip>
#include <new>

#include "engine.h"

namespace sph {

const char* stage_name(int s) {
    static const char* names[kNumStages] = {"hash",          "clear_cells",      "sort",
                                            "cell_ranges",   "reorder",          "density",
                                            "force",         "ts_state",         "ts_collect",
                                            "ts_seed_project", "ts_sub_sort",    "ts_sub_physics",
                                            "wall_mlp",    "sub_step_project"};
    return (s >= 0 && s < kNumStages) ? names[s] : "";
}

DevScene make_dev_scene(const sph_params& P, const sph_boundary& B) {
    DevScene S;
    std::memset(&S, 0, sizeof(S));
    S.ox = P.grid_origin[0]; S.oy = P.grid_origin[1]; S.oz = P.grid_origin[2];
    S.cell = P.grid_spacing;
    S.nx = P.grid_dim[0]; S.ny = P.grid_dim[1]; S.nz = P.grid_dim[2];
    S.cell_sub = P.grid_spacing_sub;
    S.snx = P.grid_dim_sub[0]; S.sny = P.grid_dim_sub[1]; S.snz = P.grid_dim_sub[2];
    S.h = P.compact_support;
    S.h2 = S.h * S.h;
    S.mass = P.particle_mass;
    S.stiff = P.gas_stiffness;
    S.rho0 = P.rest_density;
    S.mu = P.dynamic_viscosity;
    S.gravity = P.gravity;
    S.sigma = P.tension_coefficient;
    S.norm_thresh = P.norm_thresh;
    S.dt = P.time_step;
    S.poly6 = P.poly6; S.grad_poly6 = P.grad_poly6; S.lapl_poly6 = P.lapl_poly6;
    S.grad_spiky = P.grad_spiky; S.lapl_visc = P.lapl_visc;
    S.h_sub = P.compact_support_sub;
    S.h2_sub = S.h_sub * S.h_sub;
    S.mass_sub = P.sub_particle_mass;
    S.dt_sub = P.time_step_sub;
    S.poly6_sub = P.poly6_sub; S.grad_poly6_sub = P.grad_poly6_sub; S.lapl_poly6_sub = P.lapl_poly6_sub;
    S.grad_spiky_sub = P.grad_spiky_sub; S.lapl_visc_sub = P.lapl_visc_sub;
    S.box_ = B.origin[0]; S.boy_ = B.origin[1]; S.boz_ = B.origin[2];
    int dims[3] = {B.dims[0], B.dims[1], B.dims[2]};
    float scale[3], gstep[3], ghalf[3];
    for (int a = 0; a < 3; ++a) {
        // gBoundaryGridLength (cu:2626-2631) in float, as the reference
        float L = (float)(dims[a] - 1) * B.spacing;
        scale[a] = (float)((double)dims[a] / (double)L);
        gstep[a] = (float)((double)dims[a] / (double)(dims[a] - 1));
        ghalf[a] = 0.5f * (float)(dims[a] - 1);
    }
    S.bscale_x = scale[0]; S.bscale_y = scale[1]; S.bscale_z = scale[2];
    S.gstep_x = gstep[0]; S.gstep_y = gstep[1]; S.gstep_z = gstep[2];
    S.ghalf_x = ghalf[0]; S.ghalf_y = ghalf[1]; S.ghalf_z = ghalf[2];
    S.bspacing = B.spacing;
    S.r0 = B.rest_distance;
    S.bnx = dims[0]; S.bny = dims[1]; S.bnz = dims[2];
    S.restitution = P.restitution;
    S.collision = P.collision;
    for (int a = 0; a < 3; ++a) {
        S.split_lo[a] = P.split_lo[a];
        S.split_hi[a] = P.split_hi[a];
    }
    S.sub_placement = P.sub_placement;
    return S;
}

static bool params_valid(const sph_params& P, std::string& why) {
    for (int a = 0; a < 3; ++a) {
        if (P.grid_dim[a] < 1) { why = "grid_dim must be >= 1"; return false; }
        if (!std::isfinite(P.grid_origin[a])) { why = "grid_origin not finite"; return false; }
    }
    double cells = (double)P.grid_dim[0] * P.grid_dim[1] * P.grid_dim[2];
    if (cells > 2147483647.0) { why = "cell grid too large for 32-bit keys"; return false; }
    if (!(P.grid_spacing > 0.0f)) { why = "grid_spacing must be > 0"; return false; }
    if (!(P.compact_support > 0.0f)) { why = "compact_support must be > 0"; return false; }
    if (!(P.particle_mass > 0.0f)) { why = "particle_mass must be > 0"; return false; }
    if (!(P.rest_density > 0.0f)) { why = "rest_density must be > 0"; return false; }
    if (!(P.time_step > 0.0f)) { why = "time_step must be > 0"; return false; }
    return true;
}

static bool boundary_valid(const sph_boundary& B, std::string& why) {
    for (int a = 0; a < 3; ++a)
        if (B.dims[a] < 2) { why = "boundary dims must be >= 2"; return false; }
    if (!(B.spacing > 0.0f)) { why = "boundary spacing must be > 0"; return false; }
    if (!B.distance || !B.density || !B.viscosity) { why = "boundary arrays must be non-NULL"; return false; }
    double nodes = (double)B.dims[0] * B.dims[1] * B.dims[2];
    if (nodes > 4.0e9) { why = "boundary grid too large"; return false; }
    return true;
}

Engine::~Engine() {
    if (ts) twoscale_destroy(*this);
    neural_destroy(*this, false);
    drop_graph();
    for (hipEvent_t e : ev_pool) (void)hipEventDestroy(e);
    void* ptrs[] = {d_scenes, d_begins, d_tiles, grid4, gridd_alias ? nullptr : gridd, pos_a, vel_a, pos_b, vel_b, acc, oid_a, oid_b, key, skey,
                    perm, slot_of_oid, state_a, state_b, range, status, n_dev, masks, wall_mode,
                    render_pos, render_sub, surface};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    sorter.release();
    csort.release();
    if (stream) (void)hipStreamDestroy(stream);
}

sph_status Engine::fail(sph_status s, const std::string& msg) {
    err = msg;
    return s;
}

template <class T>
sph_status Engine::alloc(T** p, size_t count) {
    size_t b = sizeof(T) * (count ? count : 1);
    hipError_t e = hipMalloc(reinterpret_cast<void**>(p), b);
    if (e != hipSuccess) {
        *p = nullptr;
        (void)hipGetLastError();
        return fail(SPH_OUT_OF_MEMORY, std::string("hipMalloc failed: ") + hipGetErrorString(e));
    }
    bytes += b;
    return SPH_OK;
}

sph_status Engine::check_launch(const char* what) {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(SPH_DEVICE_ERROR, std::string(what) + ": " + hipGetErrorString(e));
    return SPH_OK;
}

sph_status Engine::init(int S, const sph_params* P, const sph_scene* scenes) {
    if (S < 1 || !P || !scenes) return fail(SPH_INVALID_ARG, "need at least one scene");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) {
        (void)hipGetLastError();
        return fail(SPH_NO_DEVICE, "no HIP device visible");
    }
    (void)hipGetDevice(&device);
    nscenes = S;
    if (S > (1 << 24)) return fail(SPH_INVALID_ARG, "at most 2^24 scenes");
    params.assign(P, P + S);
    hscenes.resize(S);
    hbegins.assign(S + 1, 0);
    cells_of.assign(S, 0);
    uint64_t ncells = 0, np = 0, nodes = 0;
    for (int s = 0; s < S; ++s) {
        std::string why;
        if (!params_valid(P[s], why)) return fail(SPH_INVALID_ARG, "scene " + std::to_string(s) + ": " + why);
        if (!boundary_valid(scenes[s].boundary, why))
            return fail(SPH_INVALID_ARG, "scene " + std::to_string(s) + ": " + why);
        if (scenes[s].num_particles < 1 || !scenes[s].positions)
            return fail(SPH_INVALID_ARG, "scene " + std::to_string(s) + ": no particles");
        for (size_t i = 0; i < 3 * (size_t)scenes[s].num_particles; ++i)
            if (!std::isfinite(scenes[s].positions[i]))
                return fail(SPH_INVALID_ARG, "scene " + std::to_string(s) + ": non-finite position");
        params[s].num_particles = scenes[s].num_particles;
        hscenes[s] = make_dev_scene(params[s], scenes[s].boundary);
        hscenes[s].cell_offset = (uint32_t)ncells;
        hscenes[s].grid_offset = nodes;
        hscenes[s].p_begin = (uint32_t)np;
        hbegins[s] = (uint32_t)np;
        cells_of[s] = (size_t)P[s].grid_dim[0] * P[s].grid_dim[1] * P[s].grid_dim[2];
        ncells += cells_of[s];
        np += scenes[s].num_particles;
        hscenes[s].p_end = (uint32_t)np;
        nodes += (uint64_t)scenes[s].boundary.dims[0] * scenes[s].boundary.dims[1] * scenes[s].boundary.dims[2];
        if (ncells > 0x7fffffffull) return fail(SPH_INVALID_ARG, "total cells exceed 32-bit keys");
        // the neighbour gathers address slots with 32-bit byte offsets
        if (np >= (1ull << 28)) return fail(SPH_INVALID_ARG, "at most 268,435,455 particles per handle or batch");
    }
    hbegins[S] = (uint32_t)np;
    N = (uint32_t)np;
    total_cells = (uint32_t)ncells;
    grid_nodes = nodes;
    key_bits = 1;
    while (key_bits < 32 && (1ull << key_bits) < (uint64_t)total_cells) ++key_bits;

    // One stream per handle.  SPH_STREAM=null puts the handle on the null
    // stream instead (for profilers that only attribute counters there).
    const char* sv = std::getenv("SPH_STREAM");
    if (sv && std::strcmp(sv, "null") == 0) {
        stream = nullptr;
    } else if (hipStreamCreateWithFlags(&stream, hipStreamNonBlocking) != hipSuccess) {
        return fail(SPH_DEVICE_ERROR, "hipStreamCreate failed");
    }
    sph_status st;
#define SPH_TRY(x) do { if ((st = (x)) != SPH_OK) return st; } while (0)
    SPH_TRY(alloc(&d_scenes, S));
    SPH_TRY(alloc(&d_begins, S + 1));
    // wavefront work list of the batch's hit-mask kernels, no entry across a
    // scene boundary: the aligned 64-slot wavefronts (each for the scene of
    // its first slot, up to that scene's end), then the slots of every scene
    // that starts inside one of them, up to the next aligned slot
    std::vector<uint2> htiles;
    if (S > 1) {
        const uint32_t n = (uint32_t)np;
        int a = 0;
        for (uint32_t f = 0; f < n; f += 64) {
            while (hbegins[a + 1] <= f) ++a;   // the scene of slot f
            const uint32_t cnt = std::min<uint32_t>(64u, hbegins[a + 1] - f);
            htiles.push_back(make_uint2(f, (uint32_t)a << 7 | cnt));
        }
        for (int s = 0; s < S; ++s) {
            const uint32_t b = hbegins[s], e = hbegins[s + 1];
            if (b == e || b % 64 == 0) continue;
            const uint32_t stop = std::min<uint32_t>(e, (b / 64 + 1) * 64);
            htiles.push_back(make_uint2(b, (uint32_t)s << 7 | (stop - b)));
        }
        ntiles = (uint32_t)htiles.size();
        SPH_TRY(alloc(&d_tiles, htiles.size()));
    }
    SPH_TRY(alloc(&grid4, nodes));
    SPH_TRY(alloc(&wall_mode, total_cells));
    SPH_TRY(alloc(&gridd, nodes));
    SPH_TRY(alloc(&pos_a, N));
    SPH_TRY(alloc(&vel_a, N));
    SPH_TRY(alloc(&pos_b, N));
    SPH_TRY(alloc(&vel_b, N));
    SPH_TRY(alloc(&acc, N));
    SPH_TRY(alloc(&oid_a, N));
    SPH_TRY(alloc(&oid_b, N));
    SPH_TRY(alloc(&key, N));
    SPH_TRY(alloc(&skey, N));
    SPH_TRY(alloc(&perm, N));
    SPH_TRY(alloc(&slot_of_oid, N));
    SPH_TRY(alloc(&state_a, N));
    SPH_TRY(alloc(&state_b, N));
    SPH_TRY(alloc(&range, total_cells));
    SPH_TRY(alloc(&status, 1));
    SPH_TRY(alloc(&n_dev, 1));
    use_cell_sort = (uint64_t)total_cells <= (uint64_t)kCellSortRatio * N + 65536;
    if (use_cell_sort) {
        if (csort.init(total_cells, N) != hipSuccess) {
            (void)hipGetLastError();
            return fail(SPH_OUT_OF_MEMORY, "cell sort buffers");
        }
        bytes += csort.bytes;
    } else {
        if (sorter.init(N) != hipSuccess) {
            (void)hipGetLastError();
            return fail(SPH_OUT_OF_MEMORY, "radix sort buffers");
        }
        bytes += sorter.bytes();
    }

    // host staging, then upload
    std::vector<float4> hp(N), hv(N), g4;
    std::vector<uint32_t> ho(N);
    for (int s = 0; s < S; ++s) {
        const sph_scene& sc = scenes[s];
        uint32_t b = hbegins[s];
        for (uint32_t i = 0; i < sc.num_particles; ++i) {
            const float* p = sc.positions + 3 * (size_t)i;
            hp[b + i] = make_float4(p[0], p[1], p[2], 0.0f);
            if (sc.velocities) {
                const float* v = sc.velocities + 3 * (size_t)i;
                hv[b + i] = make_float4(v[0], v[1], v[2], 0.0f);
            } else {
                hv[b + i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            }
            ho[b + i] = i;   // original id within its scene
        }
    }
    uint64_t off = 0;
    for (int s = 0; s < S; ++s) {
        const sph_boundary& B = scenes[s].boundary;
        size_t nn = (size_t)B.dims[0] * B.dims[1] * B.dims[2];
        g4.resize(nn);
        for (size_t q = 0; q < nn; ++q) g4[q] = make_float4(B.distance[q], B.density[q], B.viscosity[q], 0.0f);
        if (hipMemcpy(grid4 + off, g4.data(), nn * sizeof(float4), hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(gridd + off, B.distance, nn * sizeof(float), hipMemcpyHostToDevice) != hipSuccess)
            return fail(SPH_DEVICE_ERROR, "boundary upload failed");
        off += nn;
    }
    if (hipMemcpy(pos_a, hp.data(), N * sizeof(float4), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(vel_a, hv.data(), N * sizeof(float4), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(oid_a, ho.data(), N * sizeof(uint32_t), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(d_scenes, hscenes.data(), S * sizeof(DevScene), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(d_begins, hbegins.data(), (S + 1) * sizeof(uint32_t), hipMemcpyHostToDevice) != hipSuccess ||
        (ntiles && hipMemcpy(d_tiles, htiles.data(), ntiles * sizeof(uint2), hipMemcpyHostToDevice) != hipSuccess) ||
        hipMemcpy(n_dev, &N, sizeof(uint32_t), hipMemcpyHostToDevice) != hipSuccess)
        return fail(SPH_DEVICE_ERROR, "state upload failed");
    // cell tables start empty (Init, cu:1944-1946); state bytes start at 0
    if (hipMemset(range, 0xFF, (size_t)total_cells * sizeof(int2)) != hipSuccess ||
        hipMemset(status, 0, sizeof(uint32_t)) != hipSuccess || hipMemset(state_a, 0, N) != hipSuccess ||
        hipMemset(state_b, 0, N) != hipSuccess || hipMemset(acc, 0, N * sizeof(float4)) != hipSuccess ||
        hipMemset(pos_b, 0, N * sizeof(float4)) != hipSuccess ||
        hipMemset(vel_b, 0, N * sizeof(float4)) != hipSuccess ||
        hipMemset(skey, 0, N * sizeof(uint32_t)) != hipSuccess ||
        hipMemset(oid_b, 0, N * sizeof(uint32_t)) != hipSuccess)
        return fail(SPH_DEVICE_ERROR, "initial clear failed");
    // boundary source of every cell (far-field cells skip the lookups)
    for (int s = 0; s < S; ++s) launch_wall_modes(stream, hscenes[s], grid4, kWallGrid, wall_mode);
    if (hipDeviceSynchronize() != hipSuccess) return fail(SPH_DEVICE_ERROR, "device sync after create failed");
#undef SPH_TRY
    return SPH_OK;
}

StepBuffers Engine::buffers() const {
    StepBuffers b;
    b.n = N;
    b.nscenes = nscenes;
    b.scenes = d_scenes;
    b.begins = d_begins;
    b.tiles = d_tiles;
    b.ntiles = ntiles;
    b.grid4 = grid4;
    b.gridd = gridd;
    b.pos_a = pos_a; b.vel_a = vel_a; b.oid_a = oid_a; b.state_a = ts ? state_a : nullptr;
    b.pos_b = pos_b; b.vel_b = vel_b; b.oid_b = oid_b; b.state_b = ts ? state_b : nullptr;
    b.acc = acc;
    b.key = key;
    b.skey = skey;
    b.perm = perm;
    b.slot_of_oid = ts ? slot_of_oid : nullptr;
    b.range = range;
    b.status = status;
    b.masks = tuning.hit_masks ? masks : nullptr;
    b.nr.mode = wall_mode;
    b.nr.skey = skey;
    b.nr.rec = neural_records(*this);
    b.nr.need = neural_need(*this);
    b.nr.list_count = neural_list_count(*this);
    b.nr.nscenes = nscenes;
    return b;
}

hipEvent_t Engine::next_event() {
    if (ev_used == ev_pool.size()) {
        hipEvent_t e = nullptr;
        if (hipEventCreate(&e) != hipSuccess) return nullptr;
        ev_pool.push_back(e);
    }
    return ev_pool[ev_used++];
}

void Engine::mark(int stage) {
    if (!timing) return;
    hipEvent_t e = next_event();
    if (!e) return;
    (void)hipEventRecord(e, stream);
    ev_rec.emplace_back(stage, e);
}

sph_status Engine::resolve_timing() {
    if (hipStreamSynchronize(stream) != hipSuccess) return fail(SPH_DEVICE_ERROR, "timing sync failed");
    hipEvent_t prev = nullptr;
    for (auto& r : ev_rec) {
        if (r.first < 0) {
            prev = r.second;
            if (r.first == kMarkStep) ++stage_steps;
            continue;
        }
        float ms = 0.0f;
        if (prev && hipEventElapsedTime(&ms, prev, r.second) == hipSuccess) {
            stage_acc_ms[r.first] += ms;
            stage_launches[r.first] += 1;
        }
        prev = r.second;
    }
    ev_rec.clear();
    ev_used = 0;
    return SPH_OK;
}

void Engine::enqueue_step(bool two_scale) {
    StepBuffers b = buffers();
    mark(kMarkStep);
    if (!keys_valid) launch_hash(stream, b);
    mark(kStageHash);
    // the counting sort clears the previous step's cells itself
    if (cells_valid && !use_cell_sort) launch_cells_clear(stream, b);
    mark(kStageClear);
    if (use_cell_sort) {
        // the counting sort writes the cell ranges itself
        csort.sort(stream, key, N, skey, perm, range, cells_valid);
        mark(kStageSort);
    } else {
        sorter.sort(stream, key, nullptr, skey, perm, n_dev, key_bits);
        mark(kStageSort);
        launch_cells(stream, b);
        mark(kStageCells);
    }
    launch_reorder(stream, b);
    mark(kStageReorder);
    // a two-scale step's split flags come from the hit-mask density pass
    if (two_scale && b.masks) b.split = twoscale_split_refs(*this);
    launch_density(stream, b, tuning.density_lanes, tuning.block);
    mark(kStageDensity);
    if (two_scale) {
        twoscale_enqueue(*this, b.masks);
        b.state_b = twoscale_final_state(*this);
    }
    if (nb) {
        neural_enqueue(*this);   // normals where the density pass found the wall penalty active
        mark(kStageWallMlp);
    }
    launch_force(stream, b, tuning.hit_masks ? tuning.hit_list : tuning.force_lanes, tuning.block);
    mark(kStageForce);
    last_two_scale = two_scale;
    keys_valid = true;
    cells_valid = true;
}

sph_status Engine::step(int n, bool two_scale) {
    if (n < 0) return fail(SPH_INVALID_ARG, "n_steps must be >= 0");
    if (two_scale && !ts) {
        sph_status st = twoscale_create(*this);
        if (st != SPH_OK) return st;
        drop_graph();   // the step's buffers now include the sub-particle pass's
    }
    for (int i = 0; i < n; ++i) {
        if (retune_every && keys_valid && steps > 0 && steps % retune_every == 0) {
            sph_status st = autotune(2, nullptr);
            if (st != SPH_OK) return st;
        }
        // steady state: every buffer pointer and launch of the plain step is
        // fixed, so the step is captured once and replayed.  Two-scale steps
        // run eagerly: replayed from a graph they measured slower (0.250
        // against 0.215 ms at 262,144 particles; the runtime copies between
        // replays of that ~35-kernel graph and idles around the copy).
        bool graphable = use_graph && !timing && keys_valid && cells_valid && !two_scale;
        if (graphable) {
            hipGraph_t& g = graph;
            hipGraphExec_t& gx = graph_exec;
            if (!gx) {
                if (hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal) != hipSuccess)
                    return fail(SPH_DEVICE_ERROR, "hipStreamBeginCapture failed");
                enqueue_step(two_scale);
                if (hipStreamEndCapture(stream, &g) != hipSuccess)
                    return fail(SPH_DEVICE_ERROR, "hipStreamEndCapture failed");
                if (hipGraphInstantiate(&gx, g, nullptr, nullptr, 0) != hipSuccess)
                    return fail(SPH_DEVICE_ERROR, "hipGraphInstantiate failed");
            }
            if (hipGraphLaunch(gx, stream) != hipSuccess)
                return fail(SPH_DEVICE_ERROR, "hipGraphLaunch failed");
            last_two_scale = two_scale;
        } else {
            enqueue_step(two_scale);
            // bound the number of outstanding timing events
            if (timing && ev_rec.size() > 4096) {
                sph_status st = resolve_timing();
                if (st != SPH_OK) return st;
            }
        }
        sph_status st = check_launch("step");
        if (st != SPH_OK) return st;
        ++steps;
    }
    return SPH_OK;
}

sph_status Engine::sync() {
    hipError_t e = hipStreamSynchronize(stream);
    if (e != hipSuccess) return fail(SPH_DEVICE_ERROR, std::string("stream sync: ") + hipGetErrorString(e));
    uint32_t s = 0;
    if (hipMemcpy(&s, status, sizeof(uint32_t), hipMemcpyDeviceToHost) != hipSuccess)
        return fail(SPH_DEVICE_ERROR, "status read failed");
    if (s & kStatusNonFinite) return fail(SPH_NONFINITE_STATE, "non-finite particle state");
    if (s & kStatusCapacity) return fail(SPH_CAPACITY_EXCEEDED, "sub-particle capacity exceeded");
    if (s & kStatusOutOfDomain) return fail(SPH_OUT_OF_DOMAIN, "particle left the cell grid");
    return SPH_OK;
}

sph_status Engine::upload_params(int s, const sph_params& p) {
    const sph_params& o = params[s];
    for (int a = 0; a < 3; ++a)
        if (o.grid_dim[a] != p.grid_dim[a] || o.grid_origin[a] != p.grid_origin[a] ||
            o.grid_dim_sub[a] != p.grid_dim_sub[a])
            return fail(SPH_INVALID_ARG, "grid geometry may not change");
    if (o.grid_spacing != p.grid_spacing || o.grid_spacing_sub != p.grid_spacing_sub)
        return fail(SPH_INVALID_ARG, "grid spacing may not change");
    std::string why;
    if (!params_valid(p, why)) return fail(SPH_INVALID_ARG, why);
    DevScene d = hscenes[s];
    // keep boundary and layout fields, refresh the physics
    sph_boundary B;
    std::memset(&B, 0, sizeof(B));
    B.origin[0] = d.box_; B.origin[1] = d.boy_; B.origin[2] = d.boz_;
    B.spacing = d.bspacing;
    B.dims[0] = d.bnx; B.dims[1] = d.bny; B.dims[2] = d.bnz;
    B.rest_distance = d.r0;
    DevScene n = make_dev_scene(p, B);
    n.cell_offset = d.cell_offset;
    n.grid_offset = d.grid_offset;
    n.p_begin = d.p_begin;
    n.p_end = d.p_end;
    n.dist_shift = d.dist_shift;
    hscenes[s] = n;
    params[s] = p;
    params[s].num_particles = d.p_end - d.p_begin;
    if (hipMemcpyAsync(d_scenes + s, &hscenes[s], sizeof(DevScene), hipMemcpyHostToDevice, stream) != hipSuccess)
        return fail(SPH_DEVICE_ERROR, "parameter upload failed");
    if (hipStreamSynchronize(stream) != hipSuccess) return fail(SPH_DEVICE_ERROR, "parameter upload sync");
    drop_graph();
    return SPH_OK;
}

namespace {

__global__ __launch_bounds__(256) void k_distance_copy(const float4* __restrict__ grid4, uint64_t nodes,
                                                       float* __restrict__ gridd) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nodes; i += (uint64_t)gridDim.x * blockDim.x)
        gridd[i] = grid4[i].x;
}

}  // namespace

sph_status Engine::upload_scenes() {
    if (hipMemcpyAsync(d_scenes, hscenes.data(), nscenes * sizeof(DevScene), hipMemcpyHostToDevice, stream) !=
            hipSuccess ||
        hipStreamSynchronize(stream) != hipSuccess)
        return fail(SPH_DEVICE_ERROR, "scene upload failed");
    return SPH_OK;
}

// The float copy of the distance grid serves the central-difference normals
// and the collision distance (4-byte texels, a quarter of the float4 grid's
// lookup bytes).  Once every scene's normals come from a network, only the
// cells that fall back to the grid read it, and they read channel x of the
// float4 grid instead: the same values, so results do not change.
sph_status Engine::alias_distance() {
    if (gridd_alias) return SPH_OK;
    if (hipStreamSynchronize(stream) != hipSuccess) return fail(SPH_DEVICE_ERROR, "sync before freeing the distance copy");
    (void)hipFree(gridd);
    bytes -= sizeof(float) * (grid_nodes ? grid_nodes : 1);
    gridd = reinterpret_cast<float*>(grid4);
    gridd_alias = true;
    for (DevScene& d : hscenes) d.dist_shift = 2;
    drop_graph();
    return upload_scenes();
}

sph_status Engine::own_distance() {
    if (!gridd_alias) return SPH_OK;
    float* p = nullptr;
    const size_t nb = sizeof(float) * (grid_nodes ? grid_nodes : 1);
    if (hipMalloc(&p, nb) != hipSuccess) {
        (void)hipGetLastError();
        return fail(SPH_OUT_OF_MEMORY, "distance copy");
    }
    const uint64_t blocks = std::min<uint64_t>((grid_nodes + 255) / 256, 65536);
    hipLaunchKernelGGL(k_distance_copy, dim3((unsigned)std::max<uint64_t>(blocks, 1)), dim3(256), 0, stream, grid4,
                       (uint64_t)grid_nodes, p);
    if (hipGetLastError() != hipSuccess) {
        (void)hipFree(p);
        return fail(SPH_DEVICE_ERROR, "distance copy kernel");
    }
    bytes += nb;
    gridd = p;
    gridd_alias = false;
    for (DevScene& d : hscenes) d.dist_shift = 0;
    drop_graph();
    return upload_scenes();
}

void Engine::drop_graph() {
    if (graph_exec) (void)hipGraphExecDestroy(graph_exec);
    if (graph) (void)hipGraphDestroy(graph);
    graph_exec = nullptr;
    graph = nullptr;
}

static bool lanes_ok(int l) { return l == 1 || l == 2 || l == 4 || l == 8 || l == 16; }
static bool list_ok(int c) { return c == 32 || c == 40 || c == 48 || c == 64; }

sph_status Engine::set_tuning(const sph_tuning& t_in) {
    sph_tuning t = t_in;
    if (t.block == 0) t.block = 256;
    if (t.block != 128 && t.block != 256) return fail(SPH_INVALID_ARG, "block must be 128 or 256 (0 = 256)");
    if (!lanes_ok(t.density_lanes) || !lanes_ok(t.force_lanes))
        return fail(SPH_INVALID_ARG, "lanes per particle must be 1, 2, 4, 8 or 16");
    if (t.hit_masks != 0 && t.hit_masks != 1) return fail(SPH_INVALID_ARG, "hit_masks must be 0 or 1");
    if (!list_ok(t.hit_list)) return fail(SPH_INVALID_ARG, "hit_list must be 32, 40, 48 or 64");
    if (t.hit_masks && !masks) {
        sph_status st = alloc(&masks, row_words_u64(N));
        if (st != SPH_OK) return st;
        bytes += sizeof(uint64_t) * row_words_u64(N);
    }
    if (t.density_lanes != tuning.density_lanes || t.force_lanes != tuning.force_lanes ||
        t.hit_masks != tuning.hit_masks || t.hit_list != tuning.hit_list || t.block != tuning.block)
        drop_graph();
    tuning = t;
    return SPH_OK;
}

// Times the variants of the density and force kernels on the current sorted
// state (the inputs of the last step): every lanes-per-particle variant of
// each, and the hit-mask pair with every hit-list capacity.  Keeps the
// fastest density + force combination and re-runs both with it, so the state
// is what the last step would have produced with it.  Both kernels only read
// the sorted buffers, so repeating them is safe.
sph_status Engine::autotune(int reps, sph_tuning* out) {
    if (reps < 1) return fail(SPH_INVALID_ARG, "reps must be >= 1");
    if (!keys_valid) return fail(SPH_INVALID_ARG, "autotune needs the state of one step: advance once first");
    if (timing) {
        sph_status st = resolve_timing();
        if (st != SPH_OK) return st;
    }
    if (!masks) {
        sph_status st = alloc(&masks, row_words_u64(N));
        if (st != SPH_OK) return st;
        bytes += sizeof(uint64_t) * row_words_u64(N);
    }
    StepBuffers plain = buffers();
    plain.masks = nullptr;
    if (last_two_scale && ts) plain.state_b = twoscale_final_state(*this);
    StepBuffers withm = plain;
    withm.masks = masks;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
        if (e0) (void)hipEventDestroy(e0);
        return fail(SPH_DEVICE_ERROR, "event creation failed");
    }
    bool ok = true;
    auto time_it = [&](auto&& run) -> float {
        run();  // warm-up
        (void)hipEventRecord(e0, stream);
        for (int r = 0; r < reps; ++r) run();
        (void)hipEventRecord(e1, stream);
        if (hipEventSynchronize(e1) != hipSuccess) {
            ok = false;
            return 1e30f;
        }
        float ms = 0.0f;
        (void)hipEventElapsedTime(&ms, e0, e1);
        return ms;
    };
    // block sizes apply to the one-lane and hit-mask variants (the
    // cooperative ones keep 256 threads): one block size for the pair
    static const int cand[] = {1, 2, 4, 8, 16};
    static const int caps[] = {32, 40, 48, 64};
    static const int blks[] = {128, 256};
    sph_tuning pick = tuning;
    float best_plain = 1e30f, best_masks = 1e30f;
    sph_tuning pick_plain = tuning, pick_masks = tuning;
    for (int B : blks) {
        float best_d = 1e30f, best_f = 1e30f;
        int ld = 1, lf = 1;
        for (int L : cand) {
            if (L > 1 && B != 256) continue;
            float ms = time_it([&] { launch_density(stream, plain, L, B); });
            if (ms < best_d) { best_d = ms; ld = L; }
        }
        for (int L : cand) {
            if (L > 1 && B != 256) continue;
            float ms = time_it([&] { launch_force(stream, plain, L, B); });
            if (ms < best_f) { best_f = ms; lf = L; }
        }
        if (best_d + best_f < best_plain) {
            best_plain = best_d + best_f;
            pick_plain.density_lanes = ld;
            pick_plain.force_lanes = lf;
            pick_plain.block = B;
        }
        float t_dm = time_it([&] { launch_density(stream, withm, 1, B); });
        for (int c : caps) {
            float ms = time_it([&] { launch_force(stream, withm, c, B); });
            if (t_dm + ms < best_masks) {
                best_masks = t_dm + ms;
                pick_masks.hit_list = c;
                pick_masks.block = B;
            }
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (!ok) return fail(SPH_DEVICE_ERROR, "autotune: kernel failed");
    if (best_masks < best_plain) {
        pick = pick_masks;
        pick.hit_masks = 1;
    } else {
        pick = pick_plain;
        pick.hit_masks = 0;
    }
    sph_status st = set_tuning(pick);
    if (st != SPH_OK) return st;
    StepBuffers& b = tuning.hit_masks ? withm : plain;
    launch_density(stream, b, tuning.density_lanes, tuning.block);
    launch_force(stream, b, tuning.hit_masks ? tuning.hit_list : tuning.force_lanes, tuning.block);
    st = check_launch("autotune");
    if (st != SPH_OK) return st;
    if (out) *out = tuning;
    return SPH_OK;
}

sph_status Engine::read_field(int s, sph_field f, void* dst, size_t nbytes) {
    if (s < 0 || s >= nscenes) return fail(SPH_INVALID_ARG, "scene index out of range");
    if (!dst && nbytes) return fail(SPH_INVALID_ARG, "NULL destination");
    sph_status st = sync();
    if (st != SPH_OK && st != SPH_OUT_OF_DOMAIN && st != SPH_NONFINITE_STATE) return st;
    const uint32_t b = hbegins[s], n = hbegins[s + 1] - hbegins[s];
    auto need = [&](size_t want) -> sph_status {
        if (nbytes != want) {
            char buf[128];
            std::snprintf(buf, sizeof(buf), "field %d needs %zu bytes, got %zu", (int)f, want, nbytes);
            return fail(SPH_INVALID_ARG, buf);
        }
        return SPH_OK;
    };
    auto copy = [&](void* host, const void* dev, size_t bytes_) -> sph_status {
        if (hipMemcpy(host, dev, bytes_, hipMemcpyDeviceToHost) != hipSuccess)
            return fail(SPH_DEVICE_ERROR, "device-to-host copy failed");
        return SPH_OK;
    };
    std::vector<uint32_t> oid(n);
    auto load_oid = [&](const uint32_t* src) { return copy(oid.data(), src + b, n * sizeof(uint32_t)); };
    switch (f) {
    case SPH_FIELD_POSITION:
    case SPH_FIELD_VELOCITY:
    case SPH_FIELD_ACCEL: {
        if ((st = need(3 * (size_t)n * sizeof(float))) != SPH_OK) return st;
        const float4* src = f == SPH_FIELD_POSITION ? pos_a : (f == SPH_FIELD_VELOCITY ? vel_a : acc);
        std::vector<float4> tmp(n);
        if ((st = copy(tmp.data(), src + b, n * sizeof(float4))) != SPH_OK) return st;
        if ((st = load_oid(oid_a)) != SPH_OK) return st;
        float* o = (float*)dst;
        for (uint32_t i = 0; i < n; ++i) {
            uint32_t id = oid[i];
            o[3 * (size_t)id] = tmp[i].x;
            o[3 * (size_t)id + 1] = tmp[i].y;
            o[3 * (size_t)id + 2] = tmp[i].z;
        }
        return SPH_OK;
    }
    case SPH_FIELD_DENSITY:
    case SPH_FIELD_PRESSURE: {
        if ((st = need((size_t)n * sizeof(float))) != SPH_OK) return st;
        std::vector<float4> tmp(n);
        if ((st = copy(tmp.data(), vel_a + b, n * sizeof(float4))) != SPH_OK) return st;
        if ((st = load_oid(oid_a)) != SPH_OK) return st;
        float* o = (float*)dst;
        const sph_params& P = params[s];
        for (uint32_t i = 0; i < n; ++i) {
            float rho = tmp[i].w;
            o[oid[i]] = f == SPH_FIELD_DENSITY ? rho : P.gas_stiffness * (rho - P.rest_density);
        }
        return SPH_OK;
    }
    case SPH_FIELD_STATE: {
        if ((st = need(n)) != SPH_OK) return st;
        std::vector<uint8_t> tmp(n);
        if ((st = copy(tmp.data(), state_a + b, n)) != SPH_OK) return st;
        if ((st = load_oid(oid_a)) != SPH_OK) return st;
        uint8_t* o = (uint8_t*)dst;
        for (uint32_t i = 0; i < n; ++i) o[oid[i]] = tmp[i];
        return SPH_OK;
    }
    case SPH_FIELD_SORTED_KEYS: {
        if ((st = need((size_t)n * sizeof(int32_t))) != SPH_OK) return st;
        if ((st = copy(dst, skey + b, n * sizeof(uint32_t))) != SPH_OK) return st;
        int32_t* o = (int32_t*)dst;
        for (uint32_t i = 0; i < n; ++i) o[i] -= (int32_t)hscenes[s].cell_offset;
        return SPH_OK;
    }
    case SPH_FIELD_SORTED_IDS: {
        if ((st = need((size_t)n * sizeof(int32_t))) != SPH_OK) return st;
        return copy(dst, oid_b + b, n * sizeof(uint32_t));
    }
    case SPH_FIELD_CELL_START:
    case SPH_FIELD_CELL_END: {
        size_t nc = cells_of[s];
        if ((st = need(nc * sizeof(int32_t))) != SPH_OK) return st;
        std::vector<int2> tmp(nc);
        if ((st = copy(tmp.data(), range + hscenes[s].cell_offset, nc * sizeof(int2))) != SPH_OK) return st;
        int32_t* o = (int32_t*)dst;
        for (size_t c = 0; c < nc; ++c) {
            int v = f == SPH_FIELD_CELL_START ? tmp[c].x : tmp[c].y;
            o[c] = v < 0 ? -1 : v - (int)b;
        }
        return SPH_OK;
    }
    case SPH_FIELD_WALL_SOURCE: {
        if ((st = need(n)) != SPH_OK) return st;
        std::vector<uint32_t> keys(n);
        if ((st = copy(keys.data(), skey + b, n * sizeof(uint32_t))) != SPH_OK) return st;
        if ((st = load_oid(oid_b)) != SPH_OK) return st;
        uint8_t* o = (uint8_t*)dst;
        const uint32_t off = hscenes[s].cell_offset;
        std::vector<uint8_t> m(cells_of[s]);
        if ((st = copy(m.data(), wall_mode + off, cells_of[s])) != SPH_OK) return st;
        for (uint32_t i = 0; i < n; ++i) o[oid[i]] = m[keys[i] - off];
        return SPH_OK;
    }
    default:
        if (ts && s == 0) return twoscale_read(*this, f, dst, nbytes);
        return fail(SPH_INVALID_ARG, "field not available");
    }
}

sph_status Engine::extract_surface(uint8_t* flags, uint32_t* count) {
    if (nscenes != 1) return fail(SPH_INVALID_ARG, "surface extraction is per handle (one scene)");
    if (!cells_valid) return fail(SPH_INVALID_ARG, "surface extraction needs the state of one step: advance once first");
    if (!surface) {
        sph_status st = alloc(&surface, N);
        if (st != SPH_OK) return st;
        bytes += N;
    }
    launch_surface(stream, buffers(), params[0].cm_distance_thresh, params[0].n_part_thresh, surface);
    sph_status st = check_launch("surface extraction");
    if (st != SPH_OK) return st;
    if (hipStreamSynchronize(stream) != hipSuccess || hipMemcpy(flags, surface, N, hipMemcpyDeviceToHost) != hipSuccess)
        return fail(SPH_DEVICE_ERROR, "surface extraction failed");
    if (count) {
        uint32_t c = 0;
        for (uint32_t i = 0; i < N; ++i) c += flags[i];
        *count = c;
    }
    return SPH_OK;
}

// Render interop: the storage arrays hold the last step's state in cell
// order; export the positions by original id (and the sub-particles by sub
// id) into buffers a renderer can index with the id lists.
sph_status Engine::map_render(sph_render_view* out) {
    if (nscenes != 1) return fail(SPH_INVALID_ARG, "render interop is per handle (one scene)");
    if (!render_pos) {
        sph_status st = alloc(&render_pos, N);
        if (st != SPH_OK) return st;
        bytes += sizeof(float4) * (size_t)N;
    }
    std::memset(out, 0, sizeof(*out));
    launch_export_positions(stream, N, pos_a, vel_a, oid_a, render_pos);
    sph_status st = check_launch("render export");
    if (st != SPH_OK) return st;
    out->positions = reinterpret_cast<const float*>(render_pos);
    out->num_particles = N;
    if (ts && last_two_scale) {
        if (!render_sub) {
            st = alloc(&render_sub, (size_t)8 * N);
            if (st != SPH_OK) return st;
            bytes += sizeof(float4) * (size_t)8 * N;
        }
        if ((st = twoscale_export(*this, render_sub, out)) != SPH_OK) return st;
    }
    if (hipStreamSynchronize(stream) != hipSuccess) return fail(SPH_DEVICE_ERROR, "render export failed");
    return SPH_OK;
}

}  // namespace sph

using sph::Engine;

struct sph_handle {
    Engine e;
};

struct sph_batch {
    Engine e;
};

static thread_local std::string g_create_error;

namespace sph {
void set_global_error(const std::string& msg) { g_create_error = msg; }
}  // namespace sph

extern "C" {

int32_t sph_abi_version(void) { return SPH_ABI_VERSION; }

int32_t sph_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) {
        (void)hipGetLastError();
        return 0;
    }
    return n;
}

sph_status sph_create(const sph_params* params, const sph_scene* scene, sph_handle** out) {
    if (!out) return SPH_INVALID_ARG;
    *out = nullptr;
    if (!params || !scene) {
        g_create_error = "NULL params or scene";
        return SPH_INVALID_ARG;
    }
    if (scene->num_particles != params->num_particles) {
        g_create_error = "scene->num_particles differs from params->num_particles";
        return SPH_INVALID_ARG;
    }
    sph_handle* h = new (std::nothrow) sph_handle();
    if (!h) return SPH_OUT_OF_MEMORY;
    sph_status st = h->e.init(1, params, scene);
    if (st != SPH_OK) {
        g_create_error = h->e.err;
        delete h;
        return st;
    }
    g_create_error.clear();
    *out = h;
    return SPH_OK;
}

sph_status sph_destroy(sph_handle* h) {
    if (!h) return SPH_INVALID_ARG;
    (void)hipStreamSynchronize(h->e.stream);
    delete h;
    return SPH_OK;
}

sph_status sph_step(sph_handle* h, int32_t n) { return h ? h->e.step(n, false) : SPH_INVALID_ARG; }

sph_status sph_step_two_scale(sph_handle* h, int32_t n) { return h ? h->e.step(n, true) : SPH_INVALID_ARG; }

sph_status sph_step_sub_particles(sph_handle* h, int32_t n) {
    if (!h) return SPH_INVALID_ARG;
    if (n < 0) return h->e.fail(SPH_INVALID_ARG, "n_steps must be >= 0");
    for (int i = 0; i < n; ++i) {
        sph_status st = sph::twoscale_sub_step(h->e);
        if (st != SPH_OK) return st;
        // bound the number of outstanding timing events
        if (h->e.timing && h->e.ev_rec.size() > 4096 && (st = h->e.resolve_timing()) != SPH_OK) return st;
    }
    return SPH_OK;
}

sph_status sph_sync(sph_handle* h) { return h ? h->e.sync() : SPH_INVALID_ARG; }

sph_status sph_set_params(sph_handle* h, const sph_params* p) {
    if (!h || !p) return SPH_INVALID_ARG;
    return h->e.upload_params(0, *p);
}

sph_status sph_get_params(const sph_handle* h, sph_params* out) {
    if (!h || !out) return SPH_INVALID_ARG;
    *out = h->e.params[0];
    return SPH_OK;
}

sph_status sph_read_positions(sph_handle* h, float* xyz, size_t n) {
    if (!h || !xyz) return SPH_INVALID_ARG;
    if (n != h->e.N) {
        h->e.err = "n must equal the particle count";
        return SPH_INVALID_ARG;
    }
    return h->e.read_field(0, SPH_FIELD_POSITION, xyz, 3 * n * sizeof(float));
}

sph_status sph_read_field(sph_handle* h, sph_field f, void* dst, size_t bytes) {
    return h ? h->e.read_field(0, f, dst, bytes) : SPH_INVALID_ARG;
}

sph_status sph_map_positions(sph_handle* h, const float** dev_pos4, const uint32_t** dev_oid) {
    if (!h || !dev_pos4 || !dev_oid) return SPH_INVALID_ARG;
    // After a step the storage arrays hold the cell-sorted order of that
    // step; before the first step they hold the initial (id) order.
    *dev_pos4 = reinterpret_cast<const float*>(h->e.pos_a);
    *dev_oid = h->e.oid_a;
    return SPH_OK;
}

sph_status sph_map_render(sph_handle* h, sph_render_view* out) {
    if (!h || !out) return SPH_INVALID_ARG;
    return h->e.map_render(out);
}

sph_status sph_unmap_render(sph_handle* h) { return h ? SPH_OK : SPH_INVALID_ARG; }

sph_status sph_extract_surface(sph_handle* h, uint8_t* is_surface, uint32_t* num_surface) {
    if (!h || !is_surface) return SPH_INVALID_ARG;
    return h->e.extract_surface(is_surface, num_surface);
}

sph_status sph_save_info_table(sph_handle* h, const char* filename) {
    if (!h || !filename) return SPH_INVALID_ARG;
    Engine& e = h->e;
    const uint32_t n = e.hbegins[1] - e.hbegins[0];
    const size_t nc = e.cells_of[0];
    std::vector<int32_t> ids(n), keys(n), start(nc), end(nc);
    sph_status st;
    if ((st = e.read_field(0, SPH_FIELD_SORTED_IDS, ids.data(), n * sizeof(int32_t))) != SPH_OK ||
        (st = e.read_field(0, SPH_FIELD_SORTED_KEYS, keys.data(), n * sizeof(int32_t))) != SPH_OK ||
        (st = e.read_field(0, SPH_FIELD_CELL_START, start.data(), nc * sizeof(int32_t))) != SPH_OK ||
        (st = e.read_field(0, SPH_FIELD_CELL_END, end.data(), nc * sizeof(int32_t))) != SPH_OK)
        return st;
    std::ofstream f(filename);
    if (!f) return e.fail(SPH_INVALID_ARG, std::string("cannot write ") + filename);
    using std::setw;
    f << "Number of particles " << n << std::endl;
    f << setw(8) << "index" << setw(12) << " id" << setw(12) << " hash" << setw(12) << " start" << setw(12) << " end"
      << std::endl;
    for (size_t i = 0; i < nc; ++i) {
        f << setw(8) << i;
        if (i < n) f << setw(12) << ids[i] << setw(12) << keys[i];
        else f << setw(12) << "" << setw(12) << "";
        if (start[i] < 0) f << setw(12) << "";
        else f << setw(12) << start[i];
        if (end[i] < 0) f << setw(12) << "" << std::endl;
        else f << setw(12) << end[i] << std::endl;
    }
    return f ? SPH_OK : e.fail(SPH_INVALID_ARG, std::string("write failed: ") + filename);
}

sph_status sph_save_particle_info(sph_handle* h, const char* filename) {
    if (!h || !filename) return SPH_INVALID_ARG;
    Engine& e = h->e;
    const uint32_t n = e.hbegins[1] - e.hbegins[0];
    std::vector<float> x(3 * (size_t)n), rho(n), p(n), a(3 * (size_t)n);
    sph_status st;
    if ((st = e.read_field(0, SPH_FIELD_POSITION, x.data(), x.size() * sizeof(float))) != SPH_OK ||
        (st = e.read_field(0, SPH_FIELD_DENSITY, rho.data(), rho.size() * sizeof(float))) != SPH_OK ||
        (st = e.read_field(0, SPH_FIELD_PRESSURE, p.data(), p.size() * sizeof(float))) != SPH_OK ||
        (st = e.read_field(0, SPH_FIELD_ACCEL, a.data(), a.size() * sizeof(float))) != SPH_OK)
        return st;
    std::ofstream f(filename);
    if (!f) return e.fail(SPH_INVALID_ARG, std::string("cannot write ") + filename);
    const int w = 20;
    using std::setw;
    f << setw(w) << "Index" << setw(w) << "X" << setw(w) << "Y" << setw(w) << "Z" << setw(w) << "Density"
      << setw(w) << "Pressure" << setw(w) << "Acc X" << setw(w) << "Acc Y" << setw(w) << "Acc Z" << std::endl;
    for (uint32_t i = 0; i < n; ++i) {
        f << setw(w) << i << setw(w) << x[3 * (size_t)i] << setw(w) << x[3 * (size_t)i + 1] << setw(w)
          << x[3 * (size_t)i + 2] << setw(w) << rho[i] << setw(w) << p[i] << setw(w) << a[3 * (size_t)i] << setw(w)
          << a[3 * (size_t)i + 1] << setw(w) << a[3 * (size_t)i + 2] << std::endl;
    }
    return f ? SPH_OK : e.fail(SPH_INVALID_ARG, std::string("write failed: ") + filename);
}

sph_status sph_get_counts(sph_handle* h, sph_counts* out) {
    if (!h || !out) return SPH_INVALID_ARG;
    std::memset(out, 0, sizeof(*out));
    out->num_particles = h->e.N;
    out->num_default = h->e.N;
    out->num_steps = h->e.steps;
    if (h->e.ts) return sph::twoscale_counts(h->e, out);
    return SPH_OK;
}

size_t sph_num_cells(const sph_handle* h) { return h ? h->e.total_cells : 0; }

size_t sph_device_bytes(const sph_handle* h) {
    return h ? h->e.bytes + (h->e.ts ? sph::twoscale_bytes(h->e) : 0) + sph::neural_bytes(h->e) : 0;
}

const char* sph_last_error_message(const sph_handle* h) {
    return h ? h->e.err.c_str() : g_create_error.c_str();
}

static sph_status enable_timing(Engine& e, int32_t enable) {
    sph_status st = e.resolve_timing();
    e.timing = enable != 0;
    std::memset(e.stage_acc_ms, 0, sizeof(e.stage_acc_ms));
    std::memset(e.stage_launches, 0, sizeof(e.stage_launches));
    e.stage_steps = 0;
    return st;
}

static sph_status stage_times(Engine& e, float* ms, int32_t* count) {
    if (!ms || !count) return SPH_INVALID_ARG;
    sph_status st = e.resolve_timing();
    if (st != SPH_OK) return st;
    // average device time per step of each stage since timing was enabled
    uint32_t n = e.stage_steps ? e.stage_steps : 1;
    for (int s = 0; s < sph::kNumStages && s < SPH_MAX_STAGES; ++s) ms[s] = e.stage_acc_ms[s] / (float)n;
    *count = sph::kNumStages < SPH_MAX_STAGES ? sph::kNumStages : SPH_MAX_STAGES;
    return SPH_OK;
}

sph_status sph_enable_timing(sph_handle* h, int32_t enable) {
    return h ? enable_timing(h->e, enable) : SPH_INVALID_ARG;
}

sph_status sph_stage_times(sph_handle* h, float* ms, int32_t* count) {
    return h ? stage_times(h->e, ms, count) : SPH_INVALID_ARG;
}

const char* sph_stage_name(int32_t stage) { return sph::stage_name(stage); }

static sph_status time_engine(Engine& e, int32_t n, bool two_scale, float* ms) {
    if (!ms || n < 1) return SPH_INVALID_ARG;
    hipEvent_t a, b;
    if (hipEventCreate(&a) != hipSuccess || hipEventCreate(&b) != hipSuccess) return SPH_DEVICE_ERROR;
    (void)hipEventRecord(a, e.stream);
    sph_status st = e.step(n, two_scale);
    (void)hipEventRecord(b, e.stream);
    if (hipEventSynchronize(b) != hipSuccess) st = SPH_DEVICE_ERROR;
    float t = 0.0f;
    (void)hipEventElapsedTime(&t, a, b);
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    *ms = t;
    if (st != SPH_OK) return st;
    return e.sync();
}

sph_status sph_use_graph(sph_handle* h, int32_t enable) {
    if (!h) return SPH_INVALID_ARG;
    h->e.use_graph = enable != 0;
    if (!h->e.use_graph) h->e.drop_graph();
    return SPH_OK;
}

sph_status sph_time_steps(sph_handle* h, int32_t n, int32_t two_scale, float* ms) {
    return h ? time_engine(h->e, n, two_scale != 0, ms) : SPH_INVALID_ARG;
}

sph_status sph_attach_neural_boundary(sph_handle* h, sph_mlp* net, const float* val_xyzd, size_t n_val,
                                      const sph_neural_config* cfg, sph_neural_report* report) {
    if (!h || !cfg) return SPH_INVALID_ARG;
    return sph::neural_attach(h->e, 0, net, val_xyzd, n_val, *cfg, report);
}

sph_status sph_detach_neural_boundary(sph_handle* h) {
    if (!h) return SPH_INVALID_ARG;
    sph_status st = sph::neural_destroy(h->e);
    h->e.drop_graph();
    return st;
}

sph_status sph_set_autotune_interval(sph_handle* h, int32_t steps) {
    if (!h) return SPH_INVALID_ARG;
    if (steps < 0) return h->e.fail(SPH_INVALID_ARG, "interval must be >= 0");
    h->e.retune_every = (uint32_t)steps;
    return SPH_OK;
}

sph_status sph_set_tuning(sph_handle* h, const sph_tuning* t) {
    return h && t ? h->e.set_tuning(*t) : SPH_INVALID_ARG;
}

sph_status sph_get_tuning(const sph_handle* h, sph_tuning* out) {
    if (!h || !out) return SPH_INVALID_ARG;
    *out = h->e.tuning;
    return SPH_OK;
}

sph_status sph_autotune(sph_handle* h, int32_t reps, sph_tuning* out) {
    return h ? h->e.autotune(reps, out) : SPH_INVALID_ARG;
}

// ---- batched scenes ----

sph_status sph_batch_create(int32_t n_scenes, const sph_params* params, const sph_scene* scenes, sph_batch** out) {
    if (!out) return SPH_INVALID_ARG;
    *out = nullptr;
    if (n_scenes < 1 || !params || !scenes) {
        g_create_error = "need at least one scene";
        return SPH_INVALID_ARG;
    }
    sph_batch* b = new (std::nothrow) sph_batch();
    if (!b) return SPH_OUT_OF_MEMORY;
    sph_status st = b->e.init(n_scenes, params, scenes);
    if (st != SPH_OK) {
        g_create_error = b->e.err;
        delete b;
        return st;
    }
    *out = b;
    return SPH_OK;
}

sph_status sph_batch_destroy(sph_batch* b) {
    if (!b) return SPH_INVALID_ARG;
    (void)hipStreamSynchronize(b->e.stream);
    delete b;
    return SPH_OK;
}

sph_status sph_batch_step(sph_batch* b, int32_t n) { return b ? b->e.step(n, false) : SPH_INVALID_ARG; }

sph_status sph_batch_sync(sph_batch* b) { return b ? b->e.sync() : SPH_INVALID_ARG; }

sph_status sph_batch_read_field(sph_batch* b, int32_t scene, sph_field f, void* dst, size_t bytes) {
    return b ? b->e.read_field(scene, f, dst, bytes) : SPH_INVALID_ARG;
}

sph_status sph_batch_time_steps(sph_batch* b, int32_t n, float* ms) {
    return b ? time_engine(b->e, n, false, ms) : SPH_INVALID_ARG;
}

sph_status sph_batch_enable_timing(sph_batch* b, int32_t enable) {
    return b ? enable_timing(b->e, enable) : SPH_INVALID_ARG;
}

sph_status sph_batch_stage_times(sph_batch* b, float* ms, int32_t* count) {
    return b ? stage_times(b->e, ms, count) : SPH_INVALID_ARG;
}

sph_status sph_batch_use_graph(sph_batch* b, int32_t enable) {
    if (!b) return SPH_INVALID_ARG;
    b->e.use_graph = enable != 0;
    return SPH_OK;
}

sph_status sph_batch_set_tuning(sph_batch* b, const sph_tuning* t) {
    return b && t ? b->e.set_tuning(*t) : SPH_INVALID_ARG;
}

sph_status sph_batch_get_tuning(const sph_batch* b, sph_tuning* out) {
    if (!b || !out) return SPH_INVALID_ARG;
    *out = b->e.tuning;
    return SPH_OK;
}

sph_status sph_batch_autotune(sph_batch* b, int32_t reps, sph_tuning* out) {
    return b ? b->e.autotune(reps, out) : SPH_INVALID_ARG;
}

sph_status sph_batch_set_autotune_interval(sph_batch* b, int32_t steps) {
    if (!b) return SPH_INVALID_ARG;
    if (steps < 0) return b->e.fail(SPH_INVALID_ARG, "interval must be >= 0");
    b->e.retune_every = (uint32_t)steps;
    return SPH_OK;
}

sph_status sph_batch_attach_neural_boundary(sph_batch* b, int32_t scene, sph_mlp* net, const float* val_xyzd,
                                            size_t n_val, const sph_neural_config* cfg, sph_neural_report* report) {
    if (!b || !cfg) return SPH_INVALID_ARG;
    return sph::neural_attach(b->e, scene, net, val_xyzd, n_val, *cfg, report);
}

uint64_t sph_batch_num_particles(const sph_batch* b) { return b ? b->e.N : 0; }

const char* sph_batch_last_error_message(const sph_batch* b) { return b ? b->e.err.c_str() : g_create_error.c_str(); }

}  // extern "C"
