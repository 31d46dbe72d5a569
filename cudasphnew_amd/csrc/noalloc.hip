// noalloc.hip -- sph_step_noalloc: the plain step (sph_step's kernels) over a
// particle set in caller device memory, with every other buffer carved out
// of one file-scope device array.  For callers that cannot allocate (code
// under HIP graph capture, tight serving loops): the call allocates nothing,
// copies nothing between host and device and does not synchronise.
//
// The array ("scratch") is static device memory of the library's code
// object: each GPU the library runs on holds its own copy until the process
// ends, at the address hipGetSymbolAddress returns with that device current.
// Calls on one GPU share it; each call clears what must start cleared (the
// cell table, the counting sort's counts and block flags), so a call does
// not depend on the last one's contents, but calls that overlap on one GPU
// (two streams, two host threads) race on it.
#include <algorithm>
#include <mutex>
#include <string>

#include "cell_sort.h"
#include "engine.h"
#include "kernels.h"

namespace sph {

namespace {

constexpr uint32_t kNaMaxParticles = SPH_NOALLOC_MAX_PARTICLES;
constexpr uint32_t kNaMaxCells = SPH_NOALLOC_MAX_CELLS;
constexpr uint32_t kNaCountBlock = 4096;   // cells per block of the counting sort (cell_sort.hip)
constexpr uint32_t kNaMaxBlocks = (kNaMaxCells + kNaCountBlock - 1) / kNaCountBlock;
constexpr int kMaxDevices = 64;

constexpr size_t align256(size_t b) { return (b + 255) & ~size_t(255); }

// Layout of the scratch array (byte offsets, each region 256-byte aligned).
struct Layout {
    static constexpr size_t scene = 0;
    static constexpr size_t begins = scene + align256(sizeof(DevScene));
    static constexpr size_t status = begins + align256(2 * sizeof(uint32_t));
    static constexpr size_t pos_a = status + align256(sizeof(uint32_t));
    static constexpr size_t vel_a = pos_a + align256(sizeof(float4) * (size_t)kNaMaxParticles);
    static constexpr size_t pos_b = vel_a + align256(sizeof(float4) * (size_t)kNaMaxParticles);
    static constexpr size_t vel_b = pos_b + align256(sizeof(float4) * (size_t)kNaMaxParticles);
    static constexpr size_t acc = vel_b + align256(sizeof(float4) * (size_t)kNaMaxParticles);
    static constexpr size_t oid_a = acc + align256(sizeof(float4) * (size_t)kNaMaxParticles);
    static constexpr size_t oid_b = oid_a + align256(sizeof(uint32_t) * (size_t)kNaMaxParticles);
    static constexpr size_t key = oid_b + align256(sizeof(uint32_t) * (size_t)kNaMaxParticles);
    static constexpr size_t skey = key + align256(sizeof(uint32_t) * (size_t)kNaMaxParticles);
    static constexpr size_t perm = skey + align256(sizeof(uint32_t) * (size_t)kNaMaxParticles);
    static constexpr size_t rank = perm + align256(sizeof(uint32_t) * (size_t)kNaMaxParticles);
    static constexpr size_t range = rank + align256(sizeof(uint32_t) * (size_t)kNaMaxParticles);
    // counts padded to whole count blocks (the sort reads them as vectors)
    static constexpr size_t cnt = range + align256(sizeof(int2) * (size_t)kNaMaxCells);
    static constexpr size_t bsum = cnt + align256(sizeof(uint32_t) * (size_t)kNaMaxBlocks * kNaCountBlock);
    static constexpr size_t used = bsum + align256(sizeof(uint32_t) * (size_t)kNaMaxBlocks);
    static constexpr size_t bytes = used + align256(sizeof(uint32_t) * (size_t)kNaMaxBlocks);
};
static_assert(Layout::bytes == SPH_NOALLOC_SCRATCH_BYTES, "scratch size differs from the header's figure");

static __device__ __align__(256) unsigned char scratch[Layout::bytes];

// Device address of the scratch array per device, looked up once per device
// with that device current (each device holds its own copy).
std::mutex g_mu;
void* g_ptr[kMaxDevices] = {};

hipError_t scratch_address(int dev, unsigned char** out) {
    std::lock_guard<std::mutex> lock(g_mu);
    if (!g_ptr[dev]) {
        void* p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        g_ptr[dev] = p;
    }
    *out = static_cast<unsigned char*>(g_ptr[dev]);
    return hipSuccess;
}

constexpr int kNaThreads = 256;

// Per call: the scene record, the slot range, a cleared status word, an
// empty cell table, zero counts and block flags for the counting sort, and
// the caller's particles copied into storage order with original id = index.
__global__ __launch_bounds__(kNaThreads) void k_na_init(DevScene S, DevScene* __restrict__ d_scene,
                                                        uint32_t* __restrict__ begins, uint32_t* __restrict__ status,
                                                        uint32_t n, uint32_t ncells, uint32_t ncount,
                                                        uint32_t nblocks, int2* __restrict__ range,
                                                        uint32_t* __restrict__ cnt, uint32_t* __restrict__ used,
                                                        const float4* __restrict__ pos_in,
                                                        const float4* __restrict__ vel_in, float4* __restrict__ pos_a,
                                                        float4* __restrict__ vel_a, uint32_t* __restrict__ oid_a) {
    const uint32_t stride = gridDim.x * blockDim.x;
    const uint32_t t0 = blockIdx.x * blockDim.x + threadIdx.x;
    if (t0 == 0) {
        *d_scene = S;
        begins[0] = 0u;
        begins[1] = n;
        *status = 0u;
    }
    for (uint32_t i = t0; i < ncells; i += stride) range[i] = make_int2(-1, -1);
    for (uint32_t i = t0; i < ncount; i += stride) cnt[i] = 0u;
    for (uint32_t i = t0; i < nblocks; i += stride) used[i] = 0u;
    for (uint32_t i = t0; i < n; i += stride) {
        pos_a[i] = pos_in[i];
        vel_a[i] = vel_in[i];
        oid_a[i] = i;
    }
}

// Storage order back to the caller's order (original id).
__global__ __launch_bounds__(kNaThreads) void k_na_export(uint32_t n, const float4* __restrict__ pos_a,
                                                          const float4* __restrict__ vel_a,
                                                          const uint32_t* __restrict__ oid_a,
                                                          float4* __restrict__ pos_out,
                                                          float4* __restrict__ vel_out) {
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    const uint32_t id = oid_a[s];
    pos_out[id] = pos_a[s];
    vel_out[id] = vel_a[s];
}

sph_status fail(sph_status s, const char* msg) {
    set_global_error(msg);
    return s;
}

}  // namespace

}  // namespace sph

using namespace sph;

extern "C" sph_status sph_step_noalloc(const sph_params* P, const sph_device_boundary* B, uint32_t n, float* positions,
                                       float* velocities, int32_t n_steps, uint32_t* status_flags, void* stream) {
    if (!P || !B) return fail(SPH_INVALID_ARG, "NULL params or boundary");
    if (n_steps < 0) return fail(SPH_INVALID_ARG, "n_steps must be >= 0");
    if (n == 0 || !positions || !velocities) return fail(SPH_INVALID_ARG, "no particles");
    if (n > kNaMaxParticles) return fail(SPH_INVALID_ARG, "more particles than SPH_NOALLOC_MAX_PARTICLES");
    for (int a = 0; a < 3; ++a) {
        if (P->grid_dim[a] < 1) return fail(SPH_INVALID_ARG, "grid_dim must be >= 1");
        if (B->dims[a] < 2) return fail(SPH_INVALID_ARG, "boundary dims must be >= 2");
    }
    const uint64_t cells = (uint64_t)P->grid_dim[0] * P->grid_dim[1] * P->grid_dim[2];
    if (cells > kNaMaxCells) return fail(SPH_INVALID_ARG, "more cells than SPH_NOALLOC_MAX_CELLS");
    if (!(P->grid_spacing > 0.0f) || !(P->compact_support > 0.0f) || !(P->particle_mass > 0.0f) ||
        !(P->rest_density > 0.0f) || !(P->time_step > 0.0f))
        return fail(SPH_INVALID_ARG, "grid_spacing, compact_support, particle_mass, rest_density and time_step "
                                     "must be > 0");
    if (!(B->spacing > 0.0f) || !B->grid || !B->distance) return fail(SPH_INVALID_ARG, "boundary grids missing");
    if (n_steps == 0) return SPH_OK;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) {
        (void)hipGetLastError();
        return fail(SPH_NO_DEVICE, "no HIP device current");
    }
    unsigned char* base = nullptr;
    if (scratch_address(dev, &base) != hipSuccess) {
        (void)hipGetLastError();
        return fail(SPH_DEVICE_ERROR, "hipGetSymbolAddress of the scratch array failed");
    }
    hipStream_t st = static_cast<hipStream_t>(stream);
    // scene record from the parameters (make_dev_scene reads only the
    // boundary's geometry)
    sph_boundary geo{};
    for (int a = 0; a < 3; ++a) {
        geo.origin[a] = B->origin[a];
        geo.dims[a] = B->dims[a];
    }
    geo.spacing = B->spacing;
    geo.rest_distance = B->rest_distance;
    sph_params p = *P;
    p.num_particles = n;
    DevScene S = make_dev_scene(p, geo);
    S.cell_offset = 0;
    S.grid_offset = 0;
    S.p_begin = 0;
    S.p_end = n;

    StepBuffers b;
    b.n = n;
    b.nscenes = 1;
    b.scenes = reinterpret_cast<DevScene*>(base + Layout::scene);
    b.begins = reinterpret_cast<uint32_t*>(base + Layout::begins);
    b.grid4 = reinterpret_cast<const float4*>(B->grid);
    b.gridd = B->distance;
    b.pos_a = reinterpret_cast<float4*>(base + Layout::pos_a);
    b.vel_a = reinterpret_cast<float4*>(base + Layout::vel_a);
    b.pos_b = reinterpret_cast<float4*>(base + Layout::pos_b);
    b.vel_b = reinterpret_cast<float4*>(base + Layout::vel_b);
    b.acc = reinterpret_cast<float4*>(base + Layout::acc);
    b.oid_a = reinterpret_cast<uint32_t*>(base + Layout::oid_a);
    b.oid_b = reinterpret_cast<uint32_t*>(base + Layout::oid_b);
    b.key = reinterpret_cast<uint32_t*>(base + Layout::key);
    b.skey = reinterpret_cast<uint32_t*>(base + Layout::skey);
    b.perm = reinterpret_cast<uint32_t*>(base + Layout::perm);
    b.range = reinterpret_cast<int2*>(base + Layout::range);
    b.status = status_flags ? status_flags : reinterpret_cast<uint32_t*>(base + Layout::status);
    // state bytes, the id -> slot map, hit masks and the learned boundary
    // stay off: the plain step with one lane per particle (sph_step's
    // default variants), grid lookups everywhere

    CellSort cs;   // counting sort over the scratch's count array
    cs.cells = (uint32_t)cells;
    cs.blocks = (uint32_t)((cells + kNaCountBlock - 1) / kNaCountBlock);
    cs.cnt = reinterpret_cast<uint32_t*>(base + Layout::cnt);
    cs.rank = reinterpret_cast<uint32_t*>(base + Layout::rank);
    cs.bsum = reinterpret_cast<uint32_t*>(base + Layout::bsum);
    cs.used = reinterpret_cast<uint32_t*>(base + Layout::used);

    const uint32_t ncount = cs.blocks * kNaCountBlock;
    const uint32_t span = std::max(std::max((uint32_t)cells, ncount), n);
    const unsigned init_grid = std::min<uint32_t>((span + kNaThreads - 1) / kNaThreads, 4096u);
    hipLaunchKernelGGL(k_na_init, dim3(init_grid), dim3(kNaThreads), 0, st, S,
                       reinterpret_cast<DevScene*>(base + Layout::scene), reinterpret_cast<uint32_t*>(base + Layout::begins),
                       b.status, n, (uint32_t)cells, ncount, cs.blocks, b.range, cs.cnt, cs.used,
                       reinterpret_cast<const float4*>(positions), reinterpret_cast<const float4*>(velocities),
                       b.pos_a, b.vel_a, b.oid_a);
    for (int32_t i = 0; i < n_steps; ++i) {
        // the steps of Engine::enqueue_step (plain, counting sort): keys of
        // the caller's order first, later ones from the force pass
        if (i == 0) launch_hash(st, b);
        cs.sort(st, b.key, n, b.skey, b.perm, b.range, i > 0);
        launch_reorder(st, b);
        launch_density(st, b, 1);
        launch_force(st, b, 1);
    }
    hipLaunchKernelGGL(k_na_export, dim3((n + kNaThreads - 1) / kNaThreads), dim3(kNaThreads), 0, st, n, b.pos_a,
                       b.vel_a, b.oid_a, reinterpret_cast<float4*>(positions), reinterpret_cast<float4*>(velocities));
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(SPH_DEVICE_ERROR, hipGetErrorString(e));
    return SPH_OK;
}
