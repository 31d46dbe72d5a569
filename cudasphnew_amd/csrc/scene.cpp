// scene.cpp -- host-side scene construction for the SPH step.
//
// Replaces the reference's host set-up path:
//   * Example01 parameter derivation   particle_simulation.cu:1745-1817
//   * create_particle_box              particle_simulation.cu:2866-2917
//   * set_simulation_domain            particle_simulation.cu:2919-2941
//   * SphInComplexShapes grid builder  SphInComplexShapes.cpp:34-64,203-313,
//                                      366-427,562-652,666-747
// The container-box part of the grid builder reproduces the reference's FP32
// arithmetic step by step (including its double-precision kernel constants),
// so that the golden slices distances.ppm / viscosities.ppm are reproduced.
// Obstacle solids (sphere, cone, box) reuse the same ghost-particle
// construction with a lattice anchored on each solid's bounding box.
#include "sph.h"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

namespace {

constexpr double kPi = 3.14159265358979323846;  // M_PI of util.h:18

// ---------------------------------------------------------------------------
// Counter-based jitter: splitmix64 of (seed, particle, axis).
uint64_t splitmix64(uint64_t x) {
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

// Uniform in [-1, 1) with 24 random bits.
float unit_pm1(uint64_t seed, uint64_t counter) {
    uint64_t r = splitmix64(seed * 0xD1B54A32D192ED03ull ^ splitmix64(counter));
    float u = (float)(r >> 40) * (1.0f / 16777216.0f);  // [0,1)
    return 2.0f * u - 1.0f;
}

// Uniform double in [0,1) for scene randomisation (batch configs).
double unit01(uint64_t seed, uint64_t counter) {
    uint64_t r = splitmix64(seed ^ splitmix64(counter * 0x9E3779B97F4A7C15ull + 7));
    return (double)(r >> 11) * (1.0 / 9007199254740992.0);
}

// ---------------------------------------------------------------------------
// Signed distance to an axis-aligned box, reference convention: negative
// inside, positive outside.  computeSignedDistancePoint3Box3
// (SphInComplexShapes.cpp:714-747): inside, the per-axis |t|-e of smallest
// magnitude (tie rule :727-739); outside, the Euclidean point-box distance
// (WildMagic DistPoint3Box3, restated: sqrt of the sum of squared per-axis
// excesses in axis order).
float box_sdf_ref(const float p[3], const float c[3], const float e[3]) {
    float t0 = p[0] - c[0], t1 = p[1] - c[1], t2 = p[2] - c[2];
    float a = std::fabs(t0) - e[0];
    float b = std::fabs(t1) - e[1];
    float cc = std::fabs(t2) - e[2];
    float d = (std::fabs(a) < std::fabs(b)) ? a : b;
    if (std::fabs(d) > std::fabs(cc)) d = cc;
    if (a > 0.0f || b > 0.0f || cc > 0.0f) {
        float sq = 0.0f;
        if (a > 0.0f) sq += a * a;
        if (b > 0.0f) sq += b * b;
        if (cc > 0.0f) sq += cc * cc;
        d = std::sqrt(sq);
    }
    return d;
}

// isInsideBox (SphInComplexShapes.cpp:699-712): strict interior test.
bool inside_box_open(const float p[3], const float c[3], const float e[3]) {
    return !(std::fabs(p[0] - c[0]) >= e[0] || std::fabs(p[1] - c[1]) >= e[1] ||
             std::fabs(p[2] - c[2]) >= e[2]);
}

// Signed distance of one obstacle solid, positive inside the solid.
float solid_sdf(const sph_solid& s, const float p[3]) {
    switch (s.kind) {
    case SPH_SOLID_SPHERE: {
        float dx = p[0] - s.center[0], dy = p[1] - s.center[1], dz = p[2] - s.center[2];
        return s.a[0] - std::sqrt(dx * dx + dy * dy + dz * dz);
    }
    case SPH_SOLID_CONE: {
        // Solid of revolution of the triangle (0,0),(R,0),(0,H) in the
        // (radial, height) half plane; base on y = center[1].
        float R = s.a[0], H = s.a[1];
        float dx = p[0] - s.center[0], dz = p[2] - s.center[2];
        float q = std::sqrt(dx * dx + dz * dz);
        float y = p[1] - s.center[1];
        // distance to the base segment (0..R at height 0)
        float bx = std::min(std::max(q, 0.0f), R);
        float db = std::sqrt((q - bx) * (q - bx) + y * y);
        // distance to the slant segment (R,0)-(0,H)
        float ex = -R, ey = H;
        float t = ((q - R) * ex + y * ey) / (ex * ex + ey * ey);
        t = std::min(std::max(t, 0.0f), 1.0f);
        float sx = R + t * ex, sy = t * ey;
        float ds = std::sqrt((q - sx) * (q - sx) + (y - sy) * (y - sy));
        float dist = std::min(db, ds);
        bool inside = (y >= 0.0f) && (y <= H) && (q <= R * (1.0f - y / H));
        return inside ? dist : -dist;
    }
    case SPH_SOLID_BOX:
        return -box_sdf_ref(p, s.center, s.a);
    default:
        return -1e30f;
    }
}

void solid_bbox(const sph_solid& s, float lo[3], float hi[3]) {
    switch (s.kind) {
    case SPH_SOLID_SPHERE:
        for (int i = 0; i < 3; ++i) { lo[i] = s.center[i] - s.a[0]; hi[i] = s.center[i] + s.a[0]; }
        break;
    case SPH_SOLID_CONE:
        lo[0] = s.center[0] - s.a[0]; hi[0] = s.center[0] + s.a[0];
        lo[1] = s.center[1];          hi[1] = s.center[1] + s.a[1];
        lo[2] = s.center[2] - s.a[0]; hi[2] = s.center[2] + s.a[0];
        break;
    case SPH_SOLID_BOX:
        for (int i = 0; i < 3; ++i) { lo[i] = s.center[i] - s.a[i]; hi[i] = s.center[i] + s.a[i]; }
        break;
    default:
        for (int i = 0; i < 3; ++i) { lo[i] = 0; hi[i] = -1; }
    }
}

// Union of the solids: container (outside the box) and obstacles.
float scene_sdf_point(const sph_scene_desc& d, const float p[3]) {
    float c[3], e[3];
    for (int i = 0; i < 3; ++i) {
        c[i] = 0.5f * (d.box_lo[i] + d.box_hi[i]);
        e[i] = 0.5f * (d.box_hi[i] - d.box_lo[i]);
    }
    float v = box_sdf_ref(p, c, e);
    for (int k = 0; k < d.num_obstacles && k < SPH_MAX_OBSTACLES; ++k)
        v = std::max(v, solid_sdf(d.obstacles[k], p));
    return v;
}

// Kernels of the grid builder (SphInComplexShapes.cpp:666-690): the
// coefficient is evaluated in double (M_PI is a double literal) and rounded
// to float, the rest is float arithmetic.
float density_kernel(float sqDist, float h) {
    float coeff = (float)(315.0f / (64.0f * kPi * h * h * h * h * h * h * h * h * h));
    float sqh = h * h;
    float diff = sqh - sqDist;
    if (diff < 0.0f) return 0.0f;
    return coeff * diff * diff * diff;
}

float viscosity_kernel(float sqDist, float h) {
    float dist = std::sqrt(sqDist);
    if (dist > h) return 0.0f;
    return (float)(45.0f / (kPi * h * h * h * h * h * h) * (h - dist));
}

struct GridBuilder {
    float start[3];
    float spacing;
    unsigned dims[3];
    float rest_distance;
    float support;
    std::vector<unsigned> index;       // node -> list entry (0 = default)
    std::vector<float> dist, dens, visc;  // node lists, entry 0 = default

    size_t num_nodes() const { return (size_t)dims[0] * dims[1] * dims[2]; }
    size_t node_index(unsigned i, unsigned j, unsigned k) const {
        return i + (size_t)dims[0] * (j + (size_t)dims[1] * k);
    }
    void node_pos(unsigned i, unsigned j, unsigned k, float p[3]) const {
        p[0] = start[0] + spacing * i;
        p[1] = start[1] + spacing * j;
        p[2] = start[2] + spacing * k;
    }
    // computeGridCoordinate (SphInComplexShapes.cpp:562-603).
    void grid_coord(const float p[3], int c[3]) const {
        for (int a = 0; a < 3; ++a) {
            int v = (int)((p[a] - start[a]) / spacing);
            if (v < 0) v = 0;
            if (v >= (int)dims[a]) v = (int)dims[a] - 1;
            c[a] = v;
        }
    }
    // reset (SphInComplexShapes.cpp:641-652).
    void reset() {
        index.assign(num_nodes(), 0u);
        dist.assign(1, -rest_distance);
        dens.assign(1, 0.0f);
        visc.assign(1, 0.0f);
    }
    // Node distances: SetBox's first loop (:207-230), i outermost.
    template <class F>
    void set_distances(F&& sdf) {
        for (unsigned i = 0; i < dims[0]; ++i)
            for (unsigned j = 0; j < dims[1]; ++j)
                for (unsigned k = 0; k < dims[2]; ++k) {
                    float p[3];
                    node_pos(i, j, k, p);
                    float d = sdf(p);
                    if (std::fabs(d) < support) {
                        if (std::fabs(d) > rest_distance) d = d / std::fabs(d) * rest_distance;
                        dist.push_back(d);
                        index[node_index(i, j, k)] = (unsigned)(dist.size() - 1);
                    }
                }
        dens.assign(dist.size(), 0.0f);
        visc.assign(dist.size(), 0.0f);
    }
    // Splat of one ghost particle (:271-304).
    void splat(const float g[3]) {
        float lo[3] = {g[0] - support, g[1] - support, g[2] - support};
        float hi[3] = {g[0] + support, g[1] + support, g[2] + support};
        int cs[3], ce[3];
        grid_coord(lo, cs);
        grid_coord(hi, ce);
        for (int k = cs[2]; k <= ce[2]; ++k)
            for (int j = cs[1]; j <= ce[1]; ++j)
                for (int i = cs[0]; i <= ce[0]; ++i) {
                    unsigned node = index[node_index(i, j, k)];
                    if (node == 0) continue;
                    float p[3];
                    node_pos(i, j, k, p);
                    float dx = g[0] - p[0], dy = g[1] - p[1], dz = g[2] - p[2];
                    float sq = dx * dx + dy * dy + dz * dz;
                    dens[node] += density_kernel(sq, support);
                    visc[node] += viscosity_kernel(sq, support);
                }
    }
    // Final scaling (:306-311).
    void finish(float mass, float viscosity, float rest_density) {
        float vs = viscosity * mass / rest_density;
        for (size_t i = 1; i < dist.size(); ++i) {
            dens[i] *= mass;
            visc[i] *= vs;
        }
    }
    // Create*TextureData (:366-427).
    void flatten(float* d, float* w, float* v) const {
        size_t n = num_nodes();
        for (size_t q = 0; q < n; ++q) {
            unsigned e = index[q];
            if (d) d[q] = dist[e];
            if (w) w[q] = dens[e];
            if (v) v[q] = visc[e];
        }
    }
};

// Ghost lattice of the container box (SphInComplexShapes.cpp:232-260).
template <class Emit>
void container_ghosts(const float c[3], const float e[3], float support, float spacing, Emit&& emit) {
    float d = support + spacing;
    float oe[3] = {e[0] + d, e[1] + d, e[2] + d};
    float s[3] = {c[0] - oe[0], c[1] - oe[1], c[2] - oe[2]};
    unsigned iMax = (unsigned)(2 * oe[0] / spacing);
    unsigned jMax = (unsigned)(2 * oe[1] / spacing);
    unsigned kMax = (unsigned)(2 * oe[2] / spacing);
    for (unsigned k = 0; k <= kMax; ++k)
        for (unsigned j = 0; j <= jMax; ++j)
            for (unsigned i = 0; i <= iMax; ++i) {
                float p[3] = {s[0] + (float)i * spacing, s[1] + (float)j * spacing,
                              s[2] + (float)k * spacing};
                if (!inside_box_open(p, c, e)) emit(p);
            }
}

// Ghost lattice of a solid: lattice over the solid's bounding box [blo,
// bhi] grown by support + spacing; a point is a ghost if it lies in the
// solid (sdf > 0 inside) within that depth and inside the container (the
// container's own ghosts cover the region outside it).
template <class Sdf, class Emit>
void solid_ghosts(const sph_scene_desc& dsc, const float blo[3], const float bhi[3], Sdf&& sdf, float support,
                  float spacing, Emit&& emit) {
    float d = support + spacing;
    float lo[3], hi[3];
    float c[3], e[3];
    for (int i = 0; i < 3; ++i) {
        c[i] = 0.5f * (dsc.box_lo[i] + dsc.box_hi[i]);
        e[i] = 0.5f * (dsc.box_hi[i] - dsc.box_lo[i]);
    }
    unsigned n[3];
    for (int a = 0; a < 3; ++a) {
        lo[a] = blo[a] - d;
        hi[a] = bhi[a] + d;
        n[a] = (unsigned)((hi[a] - lo[a]) / spacing);
    }
    for (unsigned k = 0; k <= n[2]; ++k)
        for (unsigned j = 0; j <= n[1]; ++j)
            for (unsigned i = 0; i <= n[0]; ++i) {
                float p[3] = {lo[0] + (float)i * spacing, lo[1] + (float)j * spacing,
                              lo[2] + (float)k * spacing};
                float v = sdf(p);
                if (v >= 0.0f && v <= d && inside_box_open(p, c, e)) emit(p);
            }
}

template <class Emit>
void obstacle_ghosts(const sph_scene_desc& dsc, const sph_solid& s, float support, float spacing,
                     Emit&& emit) {
    float lo[3], hi[3];
    solid_bbox(s, lo, hi);
    solid_ghosts(dsc, lo, hi, [&](const float p[3]) { return solid_sdf(s, p); }, support, spacing, emit);
}

// Particle mass: explicit, else rho0 times the lattice cell volume.
float particle_mass_of(const sph_scene_desc& d) {
    return d.particle_mass > 0.0f ? d.particle_mass
                                  : d.rest_density * d.lattice_spacing * d.lattice_spacing * d.lattice_spacing;
}

// Ghost lattice spacing: explicit, else std::powf(m/rho0, 1/3) (cu:2598).
float ghost_spacing_of(const sph_scene_desc& d) {
    return d.ghost_spacing > 0.0f ? d.ghost_spacing : std::pow(particle_mass_of(d) / d.rest_density, 1.0f / 3.0f);
}

// A solid given by a sampled signed distance (sph_map_solid): negative
// inside at the nodes, trilinear inside a cell whose eight nodes all hold a
// distance.  Elsewhere the point is at least `band` from the surface and
// takes +-band with the side of its nearest node.  Nodes without a distance
// take the side of the node they are first reached from in a breadth-first
// walk over the six neighbours that starts from every node with one: the
// band of such nodes separates the inside from the outside, so each
// connected region of the others lies on one side.  (With no node holding
// a distance everything is outside.)  Points off the grid are outside.
struct MapSolid {
    const sph_map_solid* m = nullptr;
    std::vector<uint8_t> inside;   // per node

    size_t at(int i, int j, int k) const { return i + (size_t)m->dims[0] * (j + (size_t)m->dims[1] * k); }
    bool present(size_t q) const { return !m->present || m->present[q]; }

    void prepare(const sph_map_solid* map) {
        m = map;
        const size_t n = (size_t)m->dims[0] * m->dims[1] * m->dims[2];
        inside.assign(n, 0);
        std::vector<uint8_t> seen(n, 0);
        std::vector<size_t> queue;
        queue.reserve(n);
        for (size_t q = 0; q < n; ++q)
            if (present(q)) {
                inside[q] = m->distance[q] < 0.0f;
                seen[q] = 1;
                queue.push_back(q);
            }
        const int nx = m->dims[0], ny = m->dims[1], nz = m->dims[2];
        for (size_t h = 0; h < queue.size(); ++h) {
            const size_t q = queue[h];
            const int i = (int)(q % nx), j = (int)((q / nx) % ny), k = (int)(q / ((size_t)nx * ny));
            const int nb[6][3] = {{i - 1, j, k}, {i + 1, j, k}, {i, j - 1, k}, {i, j + 1, k}, {i, j, k - 1}, {i, j, k + 1}};
            for (const auto& c : nb) {
                if (c[0] < 0 || c[1] < 0 || c[2] < 0 || c[0] >= nx || c[1] >= ny || c[2] >= nz) continue;
                const size_t r = at(c[0], c[1], c[2]);
                if (seen[r]) continue;
                seen[r] = 1;
                inside[r] = inside[q];
                queue.push_back(r);
            }
        }
    }

    // > 0 inside the solid (scene convention).
    float sdf(const float p[3]) const {
        int i0[3];
        float w[3];
        int nearest[3];
        for (int a = 0; a < 3; ++a) {
            const float t = (p[a] - m->lo[a]) / m->dx;
            if (!(t >= 0.0f && t <= (float)(m->dims[a] - 1))) return -m->band;
            i0[a] = std::min((int)t, m->dims[a] - 2);
            w[a] = t - (float)i0[a];
            nearest[a] = std::min((int)(t + 0.5f), m->dims[a] - 1);
        }
        float c[8];
        for (int v = 0; v < 8; ++v) {
            const size_t q = at(i0[0] + (v & 1), i0[1] + ((v >> 1) & 1), i0[2] + (v >> 2));
            if (!present(q)) {
                const bool in = inside[at(nearest[0], nearest[1], nearest[2])];
                return in ? m->band : -m->band;
            }
            c[v] = m->distance[q];
        }
        auto lerp = [](float a, float b, float t) { return a + t * (b - a); };
        const float x00 = lerp(c[0], c[1], w[0]), x10 = lerp(c[2], c[3], w[0]);
        const float x01 = lerp(c[4], c[5], w[0]), x11 = lerp(c[6], c[7], w[0]);
        return -lerp(lerp(x00, x10, w[1]), lerp(x01, x11, w[1]), w[2]);
    }
};

// A map solid is usable by a scene when every point the grid builder and
// the ghost lattice look at (within support + ghost spacing of the solid)
// lies on the grid in a cell of nodes that hold distances: the grid covers
// the solid's box grown by that depth and the band exceeds it by a cell
// diagonal.
bool map_ok(const sph_map_solid& m, float support, float ghost_spacing) {
    if (!m.distance || !(m.dx > 0.0f) || !std::isfinite(m.dx) || !(m.band > 0.0f) || !std::isfinite(m.band))
        return false;
    const float depth = support + ghost_spacing;
    if (!(m.band >= depth + std::sqrt(3.0f) * m.dx)) return false;
    for (int a = 0; a < 3; ++a) {
        if (m.dims[a] < 2 || !std::isfinite(m.lo[a]) || !std::isfinite(m.solid_lo[a]) ||
            !std::isfinite(m.solid_hi[a]) || !(m.solid_lo[a] <= m.solid_hi[a]))
            return false;
        const float hi = m.lo[a] + (float)(m.dims[a] - 1) * m.dx;
        if (!(m.lo[a] <= m.solid_lo[a] - depth && hi >= m.solid_hi[a] + depth)) return false;
    }
    return (size_t)m.dims[0] * m.dims[1] * m.dims[2] <= ((size_t)1 << 31);
}

sph_status prepare_maps(const sph_scene_desc* d, const sph_map_solid* maps, int32_t num_maps,
                               std::vector<MapSolid>& out) {
    if (num_maps < 0 || (num_maps > 0 && !maps)) return SPH_INVALID_ARG;
    out.resize((size_t)num_maps);
    for (int32_t k = 0; k < num_maps; ++k) {
        if (!map_ok(maps[k], d->boundary_support, ghost_spacing_of(*d))) return SPH_INVALID_ARG;
        out[k].prepare(&maps[k]);
    }
    return SPH_OK;
}

// Union of the scene's solids and the map solids, > 0 inside a solid.
float union_sdf(const sph_scene_desc& d, const std::vector<MapSolid>& maps, const float p[3]) {
    float v = scene_sdf_point(d, p);
    for (const MapSolid& m : maps) v = std::max(v, m.sdf(p));
    return v;
}

bool finite3(const float* v) {
    return std::isfinite(v[0]) && std::isfinite(v[1]) && std::isfinite(v[2]);
}

void fill_common(sph_scene_desc* d) {
    d->gravity = 9.81f;
    d->tension = 0.0728f;       // cu:1769
    d->norm_thresh = 15.065f;   // cu:1770
    d->collision = 1;
    d->restitution = 0.0f;      // cu:1799
    d->split_lo[0] = 0.2f;      // cu:947
    d->split_hi[0] = 0.5f;
    d->split_lo[1] = d->split_lo[2] = -INFINITY;
    d->split_hi[1] = d->split_hi[2] = INFINITY;
    d->sub_placement = 0;
    d->ghost_spacing = 0.0f;
}

// Dam break in the corner of a box.  The boundary grid spans the box grown
// by r0 on every side with a single spacing (the reference has one
// gBoundaryGridSpacing, cu:43); the first lattice particle sits outside the
// wall-penalty band, which reaches r0 + 1.5 grid spacings from a wall once
// the clamped distance field is interpolated at the reference's half-texel
// shifted coordinates.
void dam_break(sph_scene_desc* d, unsigned nx, unsigned ny, unsigned nz, float spacing, float h,
               const float box_hi[3], const int bdims[3]) {
    d->lattice[0] = nx; d->lattice[1] = ny; d->lattice[2] = nz;
    d->lattice_spacing = spacing;
    d->h = h;
    d->rest_distance = h;
    d->boundary_support = h;
    d->jitter = 0.001f;
    float margin = d->rest_distance;
    float sp = 0.0f;
    for (int a = 0; a < 3; ++a) {
        d->box_lo[a] = 0.0f;
        d->box_hi[a] = box_hi[a];
        d->cell_lo[a] = -2.0f * h;
        d->cell_hi[a] = box_hi[a] + 2.0f * h;
        sp = std::max(sp, (box_hi[a] + 2.0f * margin) / (float)(bdims[a] - 1));
    }
    d->boundary_spacing = sp;
    for (int a = 0; a < 3; ++a) {
        d->boundary_dims[a] = bdims[a];
        float span = sp * (float)(bdims[a] - 1);
        d->boundary_origin[a] = 0.5f * box_hi[a] - 0.5f * span;
    }
    float off = d->rest_distance + 1.5f * sp + 2.0f * d->jitter;
    for (int a = 0; a < 3; ++a) d->lattice_origin[a] = off;
    d->rest_density = 1000.0f;
    d->particle_mass = 1000.0f * spacing * spacing * spacing;
    d->gas_stiffness = 3.0f;
    d->viscosity = 3.5f;
    d->dt = 1e-4f;
    d->dt_sub = 1e-4f;
}

}  // namespace

extern "C" {

sph_status sph_scene_config(int32_t config, int32_t index, uint64_t seed, sph_scene_desc* d) {
    if (!d) return SPH_INVALID_ARG;
    std::memset(d, 0, sizeof(*d));
    fill_common(d);
    d->seed = seed;
    switch (config) {
    case SPH_CONFIG_EXAMPLE01: {
        // Example01 (cu:1745-1817): 34^3 lattice, no jitter, dt = 0.003.
        unsigned num = (unsigned)std::pow(40000.0, 1.0 / 3.0);
        unsigned N = num * num * num;
        d->lattice[0] = d->lattice[1] = d->lattice[2] = num;
        d->lattice_origin[0] = -0.65f; d->lattice_origin[1] = -0.45f; d->lattice_origin[2] = -0.25f;
        d->lattice_spacing = 0.5f / (float)(num - 1);
        d->jitter = 0.0f;
        d->rest_density = 998.648f;
        d->particle_mass = d->rest_density * 0.5f * 0.5f * 0.5f / (float)N;
        d->gas_stiffness = 3.0f;
        d->viscosity = 3.0f;
        d->h = std::pow((float)((3.0f * 0.5f * 0.5f * 0.5f * 20) / (4.0f * kPi * N)), 1.0f / 3.0f);
        d->dt = 0.003f;
        d->dt_sub = 0.001f;
        for (int a = 0; a < 3; ++a) { d->cell_lo[a] = -1.0f; d->cell_hi[a] = 1.0f; }
        d->box_lo[0] = -0.7f; d->box_hi[0] = 0.7f;
        d->box_lo[1] = -0.5f; d->box_hi[1] = 0.5f;
        d->box_lo[2] = -0.3f; d->box_hi[2] = 0.3f;
        // SphInComplexShapes(s, e, h/4, h, h, ...) (cu:2596-2607)
        float s[3] = {-0.8f, -0.6f, -0.4f}, e[3] = {0.8f, 0.6f, 0.4f};
        d->boundary_spacing = d->h / 4.0f;
        for (int a = 0; a < 3; ++a) {
            d->boundary_origin[a] = s[a];
            d->boundary_dims[a] = (int)((unsigned)((e[a] - s[a]) / d->boundary_spacing) + 1);
        }
        d->rest_distance = d->h;
        d->boundary_support = d->h;
        d->collision = 0;
        return SPH_OK;
    }
    case SPH_CONFIG_DAM16K: {
        float box[3] = {0.64f, 0.64f, 0.64f};
        int bd[3] = {64, 64, 64};
        dam_break(d, 32, 16, 32, 0.01f, 0.02f, box, bd);
        return SPH_OK;
    }
    case SPH_CONFIG_DAM1M: {
        float box[3] = {2.56f, 1.28f, 2.56f};
        int bd[3] = {256, 128, 256};
        dam_break(d, 128, 64, 128, 0.005f, 0.01f, box, bd);
        d->num_obstacles = 2;
        d->obstacles[0].kind = SPH_SOLID_SPHERE;
        d->obstacles[0].center[0] = 1.4f; d->obstacles[0].center[1] = 0.3f; d->obstacles[0].center[2] = 0.9f;
        d->obstacles[0].a[0] = 0.3f;
        d->obstacles[1].kind = SPH_SOLID_CONE;
        d->obstacles[1].center[0] = 0.9f; d->obstacles[1].center[1] = 0.0f; d->obstacles[1].center[2] = 1.7f;
        d->obstacles[1].a[0] = 0.25f; d->obstacles[1].a[1] = 0.5f;
        return SPH_OK;
    }
    case SPH_CONFIG_TWO_SCALE: {
        float box[3] = {1.28f, 0.768f, 1.28f};
        // grid spacing h/2 over the box grown by r0 = h
        int bd[3];
        for (int a = 0; a < 3; ++a) bd[a] = (int)std::ceil((box[a] + 2.0f * 0.016f) / 0.008f) + 1;
        dam_break(d, 64, 64, 64, 0.008f, 0.016f, box, bd);
        d->num_obstacles = 1;
        sph_solid& s = d->obstacles[0];
        s.kind = SPH_SOLID_SPHERE;
        s.center[0] = 0.68f; s.center[1] = 0.1f; s.center[2] = 0.3f;
        s.a[0] = 0.1f;
        for (int a = 0; a < 3; ++a) {
            d->split_lo[a] = s.center[a] - 0.2f;
            d->split_hi[a] = s.center[a] + 0.2f;
        }
        d->sub_placement = 1;
        return SPH_OK;
    }
    case SPH_CONFIG_BATCH: {
        if (index < 0 || index >= 64) return SPH_INVALID_ARG;
        uint64_t base = (uint64_t)index * 16;
        // log-uniform particle counts, stratified over the 64 scenes: scene
        // i draws from stratum perm[i] of [0, 1) (a seeded permutation), so
        // the total stays near its expectation (about 4M) for any seed
        int perm[64];
        for (int i = 0; i < 64; ++i) perm[i] = i;
        for (int i = 63; i > 0; --i) {
            int j = (int)((double)(i + 1) * unit01(seed, 64 * 16 + (uint64_t)i));
            std::swap(perm[i], perm[j]);
        }
        double lnlo = std::log(4096.0), lnhi = std::log(262144.0);
        double u = ((double)perm[index] + unit01(seed, base + 0)) / 64.0;
        double N = std::exp(lnlo + (lnhi - lnlo) * u);
        unsigned ny = (unsigned)std::max(4.0, std::round(std::cbrt(N / 4.0)));
        unsigned nx = 2 * ny, nz = 2 * ny;
        float sp = 0.01f, h = 0.02f;
        float bx = nx * sp, by = ny * sp, bz = nz * sp;
        double aspect = 0.5 + 1.5 * unit01(seed, base + 1);
        float sa = (float)std::sqrt(aspect);
        float margin = 2.0f * (h + sp);
        float lead = h + 3.0f * sp;   // room for the wall band before the block
        // x and z twice the block's, in the drawn aspect ratio, scaled up
        // where needed so that the free corner region holds an obstacle of
        // at least 6 h (aspect kept)
        float wx = 2.0f * bx * sa, wz = 2.0f * bz / sa;
        float need = lead + 2.0f * margin + 6.0f * h;
        float grow = std::max(1.0f, std::max((bx + need) / wx, (bz + need) / wz));
        float box[3] = {wx * grow, 4.0f * by, wz * grow};
        float ext = std::max(box[0], box[2]) + 2.0f * h;
        int bd[3];
        for (int a = 0; a < 3; ++a) bd[a] = std::max(8, (int)std::ceil(127.0f * (box[a] + 2.0f * h) / ext) + 1);
        dam_break(d, nx, ny, nz, sp, h, box, bd);
        d->seed = seed * 1000003ull + (uint64_t)index;
        // obstacles in the free corner region x > bx, z > bz
        float fx0 = d->lattice_origin[0] + bx + margin, fz0 = d->lattice_origin[2] + bz + margin;
        float fx1 = box[0] - margin, fz1 = box[2] - margin;
        float room = std::min(fx1 - fx0, fz1 - fz0);
        int nobs = 1 + (int)(2.0 * unit01(seed, base + 2));
        d->num_obstacles = 0;
        if (room > 4 * h) {
            for (int o = 0; o < nobs; ++o) {
                sph_solid& s = d->obstacles[d->num_obstacles++];
                int kind = 1 + (int)(3.0 * unit01(seed, base + 3 + 4 * o));
                float r = 0.25f * room * (float)(0.5 + 0.5 * unit01(seed, base + 4 + 4 * o));
                float cx = fx0 + r + (room - 2 * r) * (float)unit01(seed, base + 5 + 4 * o);
                float cz = fz0 + r + (fz1 - fz0 - 2 * r) * (float)unit01(seed, base + 6 + 4 * o);
                s.kind = kind;
                s.center[0] = cx; s.center[2] = cz;
                if (kind == SPH_SOLID_SPHERE) { s.center[1] = r; s.a[0] = r; }
                else if (kind == SPH_SOLID_CONE) { s.center[1] = 0.0f; s.a[0] = r; s.a[1] = 1.5f * r; }
                else { s.center[1] = 0.5f * r; s.a[0] = r; s.a[1] = 0.5f * r; s.a[2] = 0.7f * r; }
            }
        }
        return SPH_OK;
    }
    default:
        return SPH_INVALID_ARG;
    }
}

sph_status sph_scene_params(const sph_scene_desc* d, sph_params* P) {
    if (!d || !P) return SPH_INVALID_ARG;
    if (!(d->h > 0.0f) || !(d->rest_density > 0.0f) || !(d->lattice_spacing > 0.0f))
        return SPH_INVALID_ARG;
    std::memset(P, 0, sizeof(*P));
    float h = d->h;
    // set_simulation_domain (cu:2919-2941)
    for (int a = 0; a < 3; ++a) {
        P->grid_origin[a] = d->cell_lo[a];
        P->grid_dim[a] = static_cast<int>((d->cell_hi[a] - d->cell_lo[a]) / h + 0.5);
        P->grid_dim_sub[a] = static_cast<int>((d->cell_hi[a] - d->cell_lo[a]) / (h / 2.0f) + 0.5);
    }
    P->grid_spacing = h;
    P->grid_spacing_sub = h / 2.0f;
    P->num_particles = d->lattice[0] * d->lattice[1] * d->lattice[2];
    P->compact_support = h;
    P->compact_support_sub = h / 2.0f;
    P->kernel_particles = 20;
    P->time_step = d->dt;
    P->time_step_sub = d->dt_sub;
    P->particle_mass = d->particle_mass > 0.0f
                           ? d->particle_mass
                           : d->rest_density * d->lattice_spacing * d->lattice_spacing * d->lattice_spacing;
    P->sub_particle_mass = P->particle_mass / 8.0f;
    P->gas_stiffness = d->gas_stiffness;
    P->rest_density = d->rest_density;
    P->dynamic_viscosity = d->viscosity;
    P->gravity = d->gravity;
    P->tension_coefficient = d->tension;
    P->norm_thresh = d->norm_thresh;
    P->fluid_volume = (float)P->num_particles * d->lattice_spacing * d->lattice_spacing * d->lattice_spacing;
    // kernel coefficients (cu:1779-1788): double expressions rounded to float
    P->poly6 = 315.0f / (64.0f * kPi * h * h * h * h * h * h * h * h * h);
    P->grad_poly6 = -945.0f / (32.0f * kPi * h * h * h * h * h * h * h * h * h);
    P->lapl_poly6 = -945.0f / (32.0f * kPi * h * h * h * h * h * h * h * h * h);
    P->grad_spiky = -45.0f / (kPi * h * h * h * h * h * h);
    P->lapl_visc = 45.0f / (kPi * h * h * h * h * h * h);
    P->poly6_sub = P->poly6 * 512.0f;
    P->grad_poly6_sub = P->grad_poly6 * 512.0f;
    P->lapl_poly6_sub = P->lapl_poly6 * 512.0f;
    P->grad_spiky_sub = P->grad_spiky * 64.0f;
    P->lapl_visc_sub = P->lapl_visc * 64.0f;
    P->cm_distance_thresh = 0.5f;   // cu:1811
    P->n_part_thresh = 20.0f;       // cu:1812
    for (int a = 0; a < 3; ++a) {
        P->box_cen[a] = 0.5f * (d->box_lo[a] + d->box_hi[a]);
        P->box_dim[a] = 0.5f * (d->box_hi[a] - d->box_lo[a]);
        P->split_lo[a] = d->split_lo[a];
        P->split_hi[a] = d->split_hi[a];
    }
    P->restitution = d->restitution;
    P->collision = d->collision;
    P->sub_placement = d->sub_placement;
    P->max_sub_particles = 0;
    return SPH_OK;
}

sph_status sph_scene_particles(const sph_scene_desc* d, float* pos, uint32_t capacity, uint32_t* n) {
    if (!d || !n) return SPH_INVALID_ARG;
    uint64_t N = (uint64_t)d->lattice[0] * d->lattice[1] * d->lattice[2];
    if (N > 0xFFFFFFFFull) return SPH_INVALID_ARG;
    *n = (uint32_t)N;
    if (!pos) return SPH_OK;
    if (capacity < N) return SPH_INVALID_ARG;
    // create_particle_box order (cu:2899-2911): x fastest, then y, then z.
    uint64_t id = 0;
    for (uint32_t k = 0; k < d->lattice[2]; ++k)
        for (uint32_t j = 0; j < d->lattice[1]; ++j)
            for (uint32_t i = 0; i < d->lattice[0]; ++i, ++id) {
                float base[3] = {d->lattice_origin[0] + i * d->lattice_spacing,
                                 d->lattice_origin[1] + j * d->lattice_spacing,
                                 d->lattice_origin[2] + k * d->lattice_spacing};
                for (int a = 0; a < 3; ++a) {
                    float jit = d->jitter > 0.0f ? d->jitter * unit_pm1(d->seed, 3 * id + a) : 0.0f;
                    pos[3 * id + a] = base[a] + jit;
                }
            }
    return SPH_OK;
}

sph_status sph_scene_sdf(const sph_scene_desc* d, const float* xyz, size_t n, float* out) {
    if (!d || (n && (!xyz || !out))) return SPH_INVALID_ARG;
    for (size_t i = 0; i < n; ++i) out[i] = scene_sdf_point(*d, xyz + 3 * i);
    return SPH_OK;
}

static sph_status sdf_samples(const sph_scene_desc* d, const std::vector<MapSolid>& maps, uint64_t seed, size_t n,
                              float near_frac, float near_width, float* xyzd) {
    if (!d || (n && !xyzd) || !(near_frac >= 0.0f && near_frac <= 1.0f) || !(near_width > 0.0f))
        return SPH_INVALID_ARG;
    for (int a = 0; a < 3; ++a)
        if (d->boundary_dims[a] < 2) return SPH_INVALID_ARG;
    float lo[3], ext[3];
    for (int a = 0; a < 3; ++a) {
        lo[a] = d->boundary_origin[a];
        ext[a] = (float)(d->boundary_dims[a] - 1) * d->boundary_spacing;
    }
    const size_t n_near = (size_t)((double)n * near_frac + 0.5);
    uint64_t counter = 0;
    auto draw = [&](float p[3]) {
        for (int a = 0; a < 3; ++a) {
            float u = 0.5f * (unit_pm1(seed, counter++) + 1.0f);   // [0, 1)
            p[a] = lo[a] + u * ext[a];
        }
    };
    const uint64_t max_tries = 100000ull * (uint64_t)(n_near + 1);
    uint64_t tries = 0;
    for (size_t i = 0; i < n; ++i) {
        float p[3], v;
        for (;;) {
            draw(p);
            v = union_sdf(*d, maps, p);
            if (i >= n_near || std::fabs(v) < near_width) break;
            if (++tries > max_tries) return SPH_INVALID_ARG;   // band too thin to sample
        }
        xyzd[4 * i] = p[0];
        xyzd[4 * i + 1] = p[1];
        xyzd[4 * i + 2] = p[2];
        xyzd[4 * i + 3] = v;
    }
    return SPH_OK;
}

// Distance texture lookup at a world point with the arithmetic of the
// device lookup (normalized coordinate u = (x - o)/L sampled at texel
// u*dims - 0.5, clamp addressing, linear filter as a + t (b - a)).
static float distance_texture(const float* g, const int dims[3], const float o[3], float spacing,
                              const float p[3]) {
    int i0[3], i1[3];
    float w[3];
    for (int k = 0; k < 3; ++k) {
        float L = (float)(dims[k] - 1) * spacing;
        float scale = (float)((double)dims[k] / (double)L);
        float t = (p[k] - o[k]) * scale - 0.5f;
        t = std::fmin(std::fmax(t, -1.0f), (float)dims[k]);
        float fl = std::floor(t);
        w[k] = t - fl;
        int b = (int)fl;
        i0[k] = std::min(std::max(b, 0), dims[k] - 1);
        i1[k] = std::min(std::max(b + 1, 0), dims[k] - 1);
    }
    auto at = [&](int x, int y, int z) { return g[x + (size_t)dims[0] * (y + (size_t)dims[1] * z)]; };
    auto lerp = [](float a, float b, float t) { return a + t * (b - a); };
    float c00 = lerp(at(i0[0], i0[1], i0[2]), at(i1[0], i0[1], i0[2]), w[0]);
    float c10 = lerp(at(i0[0], i1[1], i0[2]), at(i1[0], i1[1], i0[2]), w[0]);
    float c01 = lerp(at(i0[0], i0[1], i1[2]), at(i1[0], i0[1], i1[2]), w[0]);
    float c11 = lerp(at(i0[0], i1[1], i1[2]), at(i1[0], i1[1], i1[2]), w[0]);
    return lerp(lerp(c00, c10, w[1]), lerp(c01, c11, w[1]), w[2]);
}

static sph_status grid_samples(const sph_scene_desc* d, const std::vector<MapSolid>& maps, uint64_t seed, size_t n,
                               float near_frac, float near_width, float* xyzd) {
    sph_status st = sdf_samples(d, maps, seed, n, near_frac, near_width, xyzd);
    if (st != SPH_OK || n == 0) return st;
    const int dims[3] = {d->boundary_dims[0], d->boundary_dims[1], d->boundary_dims[2]};
    std::vector<float> g((size_t)dims[0] * dims[1] * dims[2]);
    GridBuilder b;
    for (int a = 0; a < 3; ++a) {
        b.start[a] = d->boundary_origin[a];
        b.dims[a] = (unsigned)dims[a];
    }
    b.spacing = d->boundary_spacing;
    b.rest_distance = d->rest_distance;
    b.support = d->boundary_support;
    b.reset();
    b.set_distances([&](const float p[3]) { return union_sdf(*d, maps, p); });
    b.flatten(g.data(), nullptr, nullptr);
    for (size_t i = 0; i < n; ++i) xyzd[4 * i + 3] = distance_texture(g.data(), dims, d->boundary_origin,
                                                                      d->boundary_spacing, xyzd + 4 * i);
    return SPH_OK;
}

sph_status sph_scene_sdf_samples(const sph_scene_desc* d, uint64_t seed, size_t n, float near_frac,
                                 float near_width, float* xyzd) {
    return sdf_samples(d, {}, seed, n, near_frac, near_width, xyzd);
}

sph_status sph_scene_grid_samples(const sph_scene_desc* d, uint64_t seed, size_t n, float near_frac,
                                  float near_width, float* xyzd) {
    return grid_samples(d, {}, seed, n, near_frac, near_width, xyzd);
}

sph_status sph_scene_sdf_samples_maps(const sph_scene_desc* d, const sph_map_solid* maps, int32_t num_maps,
                                      uint64_t seed, size_t n, float near_frac, float near_width, float* xyzd) {
    if (!d) return SPH_INVALID_ARG;
    std::vector<MapSolid> ms;
    sph_status st = prepare_maps(d, maps, num_maps, ms);
    if (st != SPH_OK) return st;
    return sdf_samples(d, ms, seed, n, near_frac, near_width, xyzd);
}

sph_status sph_scene_grid_samples_maps(const sph_scene_desc* d, const sph_map_solid* maps, int32_t num_maps,
                                       uint64_t seed, size_t n, float near_frac, float near_width, float* xyzd) {
    if (!d) return SPH_INVALID_ARG;
    std::vector<MapSolid> ms;
    sph_status st = prepare_maps(d, maps, num_maps, ms);
    if (st != SPH_OK) return st;
    return grid_samples(d, ms, seed, n, near_frac, near_width, xyzd);
}

// The three grids of a scene whose solids are the container, the analytic
// obstacles and `maps` (prepared map solids).
static sph_status build_boundary(const sph_scene_desc* d, const std::vector<MapSolid>& maps, float* distance,
                                 float* density, float* viscosity) {
    GridBuilder g;
    for (int a = 0; a < 3; ++a) {
        g.start[a] = d->boundary_origin[a];
        g.dims[a] = (unsigned)d->boundary_dims[a];
    }
    g.spacing = d->boundary_spacing;
    g.rest_distance = d->rest_distance;
    g.support = d->boundary_support;
    g.reset();
    g.set_distances([&](const float p[3]) { return union_sdf(*d, maps, p); });
    const float mass = particle_mass_of(*d);
    const float gsp = ghost_spacing_of(*d);
    float c[3], e[3];
    for (int a = 0; a < 3; ++a) {
        c[a] = 0.5f * (d->box_lo[a] + d->box_hi[a]);
        e[a] = 0.5f * (d->box_hi[a] - d->box_lo[a]);
    }
    auto emit = [&](const float p[3]) { g.splat(p); };
    container_ghosts(c, e, g.support, gsp, emit);
    for (int k = 0; k < d->num_obstacles && k < SPH_MAX_OBSTACLES; ++k)
        obstacle_ghosts(*d, d->obstacles[k], g.support, gsp, emit);
    for (const MapSolid& m : maps)
        solid_ghosts(*d, m.m->solid_lo, m.m->solid_hi, [&](const float p[3]) { return m.sdf(p); }, g.support, gsp,
                     emit);
    g.finish(mass, d->viscosity, d->rest_density);
    g.flatten(distance, density, viscosity);
    return SPH_OK;
}

static bool boundary_desc_ok(const sph_scene_desc* d) {
    if (!d) return false;
    for (int a = 0; a < 3; ++a)
        if (d->boundary_dims[a] < 2) return false;
    return d->boundary_spacing > 0.0f && finite3(d->boundary_origin);
}

sph_status sph_scene_boundary(const sph_scene_desc* d, float* distance, float* density, float* viscosity) {
    if (!boundary_desc_ok(d)) return SPH_INVALID_ARG;
    return build_boundary(d, {}, distance, density, viscosity);
}

sph_status sph_scene_boundary_maps(const sph_scene_desc* d, const sph_map_solid* maps, int32_t num_maps,
                                   float* distance, float* density, float* viscosity) {
    if (!boundary_desc_ok(d)) return SPH_INVALID_ARG;
    std::vector<MapSolid> ms;
    sph_status st = prepare_maps(d, maps, num_maps, ms);
    if (st != SPH_OK) return st;
    return build_boundary(d, ms, distance, density, viscosity);
}

sph_status sph_scene_sdf_maps(const sph_scene_desc* d, const sph_map_solid* maps, int32_t num_maps,
                              const float* xyz, size_t n, float* out) {
    if (!d || (n && (!xyz || !out))) return SPH_INVALID_ARG;
    std::vector<MapSolid> ms;
    sph_status st = prepare_maps(d, maps, num_maps, ms);
    if (st != SPH_OK) return st;
    for (size_t i = 0; i < n; ++i) out[i] = union_sdf(*d, ms, xyz + 3 * i);
    return SPH_OK;
}

sph_status sph_example01_boundary(int32_t dims[3], float origin[3], float* spacing, float* rest_distance,
                                  float* distance, float* density, float* viscosity) {
    sph_scene_desc d;
    sph_status st = sph_scene_config(SPH_CONFIG_EXAMPLE01, 0, 0, &d);
    if (st != SPH_OK) return st;
    if (dims)
        for (int a = 0; a < 3; ++a) dims[a] = d.boundary_dims[a];
    if (origin)
        for (int a = 0; a < 3; ++a) origin[a] = d.boundary_origin[a];
    if (spacing) *spacing = d.boundary_spacing;
    if (rest_distance) *rest_distance = d.rest_distance;
    if (!distance && !density && !viscosity) return SPH_OK;
    // The particle spacing of the ghost lattice is std::powf(m/rho0, 1/3)
    // (cu:2598); Example01 has no obstacles, so this is SetBox exactly.
    return sph_scene_boundary(&d, distance, density, viscosity);
}

}  // extern "C"
