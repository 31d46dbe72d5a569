// sph_oracle.cpp -- CPU oracle of the SPH step.  TEST INFRASTRUCTURE ONLY:
// the checker of the HIP library (tests/, __graft_entry__.smoke and the
// cpu_baseline leg of bench.py load it); the product never does.
//
// Restates the reference step of CUDASPHNEW/particle_simulation.cu on the
// CPU, templated on the scalar type (double = parity oracle, float = CPU
// baseline transcription).  Storage is kept in original-id order and is
// never permuted, exactly as in the reference; the sort produces the id
// list (thrust::sort_by_key of (hash, id), cu:2338-2344).
//
// Definitions this restatement fixes where the reference is undefined:
//   * colLapl starts at 0 (uninitialised in cu:381 and cu:522);
//   * a particle with a coordinate (int)((x-o)/d) < 0 is given cell
//     coordinate 0 instead of a negative hash (cu:110-115), and >= dim wraps
//     as the reference's "% dim"; both raise ORC_FLAG_OUT_OF_DOMAIN;
//   * find_boundary_particles' racy byte OR (cu:1010) is restated as the
//     race-free pull: a particle becomes "boundary" if any particle within h
//     had its split bit set this step;
//   * project_quantities with no coarse neighbour (0/0 in cu:620-621) gives
//     rest density and zero pressure;
//   * boundary sub-particles, whose acceleration the reference never writes
//     (stale memory), get zero acceleration;
//   * collision response (option): collision_handling (cu:742-834) driven by
//     the boundary distance grid: the contact normal is the normalised
//     central-difference gradient, depth the interpolated distance, and the
//     restitution term is skipped when |v| = 0 (0/0 in cu:819-820).
#include "sph_oracle.h"

#include <algorithm>
#include <chrono>
#include <climits>
#include <cmath>
#include <cstddef>
#include <cstring>
#include <memory>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

constexpr double kPi = 3.14159265358979323846;

// ---------------------------------------------------------------------------
// Cell coordinates.

// compute_particle_hash's per-axis coordinate (cu:101-115):
// (int)((x - o)/d) % dim, with the out-of-domain definitions above.
template <class R>
int hash_coord(R x, float o, float d, int dim, uint32_t& flags) {
    R f = (x - (R)o) / (R)d;
    if (!(f == f)) { flags |= ORC_FLAG_NONFINITE; return 0; }
    if (f <= (R)-1) { flags |= ORC_FLAG_OUT_OF_DOMAIN; return 0; }
    if (f >= (R)2147483520.0) { flags |= ORC_FLAG_OUT_OF_DOMAIN; return 0; }
    int c = (int)f;
    if (c >= dim) { flags |= ORC_FLAG_OUT_OF_DOMAIN; c %= dim; }
    return c;
}

// compute_grid_coordinate (cu:1238-1258): (unsigned)((x + d - o)/s) % dim,
// clamped; float->unsigned saturates at 0 for negative/NaN values.
template <class R>
int range_coord(R x, R d, float o, float s, int dim) {
    R f = ((x + d) - (R)o) / (R)s;
    if (!(f > (R)0)) return 0;
    if (f >= (R)2147483520.0) return 0;
    int c = (int)(unsigned)f;
    c = c % dim;
    return std::min(std::max(c, 0), dim - 1);
}

template <class R>
int32_t cell_key(const sph_params& P, const R* x, uint32_t& flags) {
    int cx = hash_coord<R>(x[0], P.grid_origin[0], P.grid_spacing, P.grid_dim[0], flags);
    int cy = hash_coord<R>(x[1], P.grid_origin[1], P.grid_spacing, P.grid_dim[1], flags);
    int cz = hash_coord<R>(x[2], P.grid_origin[2], P.grid_spacing, P.grid_dim[2], flags);
    return P.grid_dim[0] * (P.grid_dim[1] * cz + cy) + cx;
}

template <class R>
int32_t sub_cell_key(const sph_params& P, const R* x, uint32_t& flags) {
    int cx = hash_coord<R>(x[0], P.grid_origin[0], P.grid_spacing_sub, P.grid_dim_sub[0], flags);
    int cy = hash_coord<R>(x[1], P.grid_origin[1], P.grid_spacing_sub, P.grid_dim_sub[1], flags);
    int cz = hash_coord<R>(x[2], P.grid_origin[2], P.grid_spacing_sub, P.grid_dim_sub[2], flags);
    return P.grid_dim_sub[0] * (P.grid_dim_sub[1] * cz + cy) + cx;
}

struct Range { int lo[3], hi[3]; };

template <class R>
Range neighbour_range(const float o[3], float s, const int dim[3], const R* x, R h) {
    Range r;
    for (int a = 0; a < 3; ++a) {
        r.lo[a] = range_coord<R>(x[a], -h, o[a], s, dim[a]);
        r.hi[a] = range_coord<R>(x[a], h, o[a], s, dim[a]);
    }
    return r;
}

// ---------------------------------------------------------------------------
// Sort + cell ranges: thrust::sort_by_key (stable radix sort of the hash
// with the identity id list, cu:2338-2344) and compute_cell_start_end
// (cu:155-193).  A stable counting sort gives the same order.
void sort_and_cells(uint32_t n, const int32_t* keys, size_t ncells, int32_t* sorted_keys,
                    int32_t* sorted_ids, int32_t* start, int32_t* end) {
    std::vector<uint32_t> count(ncells + 1, 0u);
    for (uint32_t i = 0; i < n; ++i) count[(size_t)keys[i] + 1]++;
    for (size_t c = 0; c < ncells; ++c) count[c + 1] += count[c];
    for (size_t c = 0; c < ncells; ++c) {
        if (count[c + 1] > count[c]) {
            start[c] = (int32_t)count[c];
            end[c] = (int32_t)count[c + 1];
        } else {
            start[c] = -1;
            end[c] = -1;
        }
    }
    std::vector<uint32_t> pos(count.begin(), count.end() - 1);
    for (uint32_t i = 0; i < n; ++i) {
        uint32_t p = pos[keys[i]]++;
        sorted_ids[p] = (int32_t)i;
        sorted_keys[p] = keys[i];
    }
}

// ---------------------------------------------------------------------------
// Boundary grids as 3D textures (set_up_3d_float_texture, cu:2959-3011:
// normalized coordinates, linear filtering, clamp addressing).  Texel
// coordinate t = u*dims - 0.5, weights frac(t), indices clamped.
struct Grid {
    float origin[3];
    float spacing;
    int dims[3];
    float r0;
    double len[3];
    std::vector<float> field[3];  // distance, density, viscosity

    void init(const sph_boundary& b) {
        for (int a = 0; a < 3; ++a) {
            origin[a] = b.origin[a];
            dims[a] = b.dims[a];
        }
        spacing = b.spacing;
        r0 = b.rest_distance;
        size_t n = (size_t)dims[0] * dims[1] * dims[2];
        const float* src[3] = {b.distance, b.density, b.viscosity};
        for (int f = 0; f < 3; ++f) field[f].assign(src[f], src[f] + n);
        // gBoundaryGridLength = (dims - 1) * spacing (cu:2626-2631), float
        for (int a = 0; a < 3; ++a) len[a] = (double)((float)(dims[a] - 1) * spacing);
    }
};

template <class R>
R tex3d(const Grid& g, int which, R u, R v, R w) {
    const float* T = g.field[which].data();
    R t[3] = {u * (R)g.dims[0] - (R)0.5, v * (R)g.dims[1] - (R)0.5, w * (R)g.dims[2] - (R)0.5};
    int i0[3], i1[3];
    R al[3];
    for (int a = 0; a < 3; ++a) {
        R fl = std::floor(t[a]);
        al[a] = t[a] - fl;
        long long b = (long long)fl;
        long long lo = std::min<long long>(std::max<long long>(b, 0), g.dims[a] - 1);
        long long hi = std::min<long long>(std::max<long long>(b + 1, 0), g.dims[a] - 1);
        i0[a] = (int)lo;
        i1[a] = (int)hi;
    }
    auto at = [&](int i, int j, int k) -> R {
        return (R)T[(size_t)i + (size_t)g.dims[0] * ((size_t)j + (size_t)g.dims[1] * k)];
    };
    // a + t (b - a): exact for a constant field (see the wall switch)
    auto lerp = [](R a, R b, R t) { return a + t * (b - a); };
    R c00 = lerp(at(i0[0], i0[1], i0[2]), at(i1[0], i0[1], i0[2]), al[0]);
    R c10 = lerp(at(i0[0], i1[1], i0[2]), at(i1[0], i1[1], i0[2]), al[0]);
    R c01 = lerp(at(i0[0], i0[1], i1[2]), at(i1[0], i0[1], i1[2]), al[0]);
    R c11 = lerp(at(i0[0], i1[1], i1[2]), at(i1[0], i1[1], i1[2]), al[0]);
    return lerp(lerp(c00, c10, al[1]), lerp(c01, c11, al[1]), al[2]);
}

// Normalized texture coordinates of a world point (cu:243-245).
template <class R>
void tex_coords(const Grid& g, const R* x, R uvw[3]) {
    for (int a = 0; a < 3; ++a) uvw[a] = (x[a] - (R)g.origin[a]) / (R)g.len[a];
}

// ---------------------------------------------------------------------------
// Simulation state.

template <class R>
struct Sim {
    sph_params P;
    Grid G;
    int threads = 1;
    uint32_t N = 0;
    uint32_t flags = 0;
    uint32_t steps = 0;
    double last_seconds = 0.0;
    size_t ncells = 0, nsubcells = 0;

    std::vector<R> x, v, a, rho, p;          // 3N, 3N, 3N, N, N
    // branch variables of the acceleration (colour-gradient norm against
    // normThresh, wall distance against r0), for tests that must skip
    // particles sitting on a discontinuity of the reference's model
    std::vector<double> cgn_out, dwall_out, sub_cgn_out;
    std::vector<uint8_t> state;              // N
    std::vector<int32_t> keys, skeys, sids, cstart, cend;

    // two-scale
    uint32_t nsplit = 0, nboundary = 0, ndefault = 0, nsub = 0;
    std::vector<int32_t> split_ids, boundary_ids, default_ids, sub_ids;
    std::vector<R> sx, sv, sa, srho, sp;     // sub-list order
    std::vector<int32_t> skeys_sub, sorted_sub, sub_skeys, sub_start, sub_end;

    void init(const sph_params& prm, const sph_boundary& b, uint32_t n, const R* pos, const R* vel,
              int nthreads) {
        P = prm;
        P.num_particles = n;
        G.init(b);
        threads = std::max(1, nthreads);
        N = n;
        x.assign(pos, pos + 3 * (size_t)n);
        v.assign(3 * (size_t)n, (R)0);
        if (vel) v.assign(vel, vel + 3 * (size_t)n);
        a.assign(3 * (size_t)n, (R)0);
        cgn_out.assign(n, 0.0);
        dwall_out.assign(n, 0.0);
        rho.assign(n, (R)0);
        p.assign(n, (R)0);
        state.assign(n, 0);
        ncells = (size_t)P.grid_dim[0] * P.grid_dim[1] * P.grid_dim[2];
        nsubcells = (size_t)P.grid_dim_sub[0] * P.grid_dim_sub[1] * P.grid_dim_sub[2];
        keys.assign(n, 0);
        skeys.assign(n, 0);
        sids.assign(n, 0);
        cstart.assign(ncells, -1);
        cend.assign(ncells, -1);
    }

    // ---- hash + sort + cell ranges (cu:2320-2369) ----
    void hash_sort_cells() {
        uint32_t f = 0;
        for (uint32_t i = 0; i < N; ++i) keys[i] = cell_key<R>(P, &x[3 * (size_t)i], f);
        flags |= f;
        sort_and_cells(N, keys.data(), ncells, skeys.data(), sids.data(), cstart.data(), cend.data());
    }

    template <class F>
    void for_neighbours(const R* xi, R hsearch, F&& fn) const {
        Range r = neighbour_range<R>(P.grid_origin, P.grid_spacing, P.grid_dim, xi, hsearch);
        for (int k = r.lo[2]; k <= r.hi[2]; ++k)
            for (int j = r.lo[1]; j <= r.hi[1]; ++j)
                for (int i = r.lo[0]; i <= r.hi[0]; ++i) {
                    int hash = P.grid_dim[0] * (P.grid_dim[1] * k + j) + i;
                    int s = cstart[hash], e = cend[hash];
                    for (int q = s; q < e; ++q) fn(sids[q]);
                }
    }

    // ---- compute_particle_density_pressure (cu:196-256, 1385-1419) ----
    void density() {
        const R h = (R)P.compact_support;
        const R poly6 = (R)P.poly6;
#pragma omp parallel for schedule(static) num_threads(threads)
        for (long long idx = 0; idx < (long long)N; ++idx) {
            int id = sids[idx];
            const R* xi = &x[3 * (size_t)id];
            R dens = 0;
            for_neighbours(xi, h, [&](int j) {
                const R* xj = &x[3 * (size_t)j];
                R dx = xj[0] - xi[0], dy = xj[1] - xi[1], dz = xj[2] - xi[2];
                R r = std::sqrt(dx * dx + dy * dy + dz * dz);
                if (r <= h) {
                    R d = h * h - r * r;
                    dens += poly6 * d * d * d;
                }
            });
            dens *= (R)P.particle_mass;
            R uvw[3];
            tex_coords(G, xi, uvw);
            dens += tex3d<R>(G, 1, uvw[0], uvw[1], uvw[2]);
            rho[id] = dens;
            p[id] = (R)P.gas_stiffness * (dens - (R)P.rest_density);
        }
    }

    // ---- compute_particle_acceleration_ifsurf (cu:315-474, 1459-1556) ----
    void acceleration() {
        const R h = (R)P.compact_support;
        const R grad = (R)P.grad_spiky, lapl = (R)P.lapl_visc;
        const R grad2 = (R)P.grad_poly6, lapl2 = (R)P.lapl_poly6;
        const R m = (R)P.particle_mass, mu = (R)P.dynamic_viscosity;
        const R dt = (R)P.time_step;
#pragma omp parallel for schedule(static) num_threads(threads)
        for (long long idx = 0; idx < (long long)N; ++idx) {
            int id = sids[idx];
            const R* xi = &x[3 * (size_t)id];
            const R* vi = &v[3 * (size_t)id];
            R rhoi = rho[id], pi = p[id], rhoi2 = rhoi * rhoi;
            R fp[3] = {0, 0, 0}, fv[3] = {0, 0, 0}, cg[3] = {0, 0, 0}, cl = 0;
            for_neighbours(xi, h, [&](int j) {
                const R* xj = &x[3 * (size_t)j];
                const R* vj = &v[3 * (size_t)j];
                R rhoj = rho[j], pj = p[j];
                R r[3] = {xi[0] - xj[0], xi[1] - xj[1], xi[2] - xj[2]};
                R rn = std::sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
                if (rn <= h && rn > (R)0) {
                    R pressure = rhoi * (pi / rhoi2 + pj / (rhoj * rhoj)) * m;
                    R d = (h - rn) * (h - rn);
                    for (int c = 0; c < 3; ++c) fp[c] -= pressure * grad * d / rn * r[c];
                    d = h - rn;
                    for (int c = 0; c < 3; ++c) fv[c] += mu * (vj[c] - vi[c]) * m / rhoj * lapl * d;
                    d = (h * h - rn * rn) * (h * h - rn * rn);
                    for (int c = 0; c < 3; ++c) cg[c] += m / rhoj * grad2 * d * r[c];
                    d = (h * h - rn * rn) * ((R)3 * h * h - (R)7 * rn * rn);
                    cl += m / rhoj * lapl2 * d;
                }
            });
            R F[3] = {fv[0] + fp[0], fv[1] + fp[1], fv[2] + fp[2]};
            R uvw[3];
            tex_coords(G, xi, uvw);
            R distWall = -tex3d<R>(G, 0, uvw[0], uvw[1], uvw[2]);
            R r0 = (R)G.r0;
            dwall_out[id] = (double)distWall;
            cgn_out[id] = std::sqrt((double)cg[0] * cg[0] + (double)cg[1] * cg[1] + (double)cg[2] * cg[2]);
            if (distWall < r0) {
                R coeff = rhoi / (dt * dt) * (r0 - distWall);
                R gn[3];
                for (int c = 0; c < 3; ++c) {
                    R dX = (R)G.spacing / (R)G.len[c];
                    R up[3] = {uvw[0], uvw[1], uvw[2]}, dn[3] = {uvw[0], uvw[1], uvw[2]};
                    up[c] += dX;
                    dn[c] -= dX;
                    gn[c] = (tex3d<R>(G, 0, up[0], up[1], up[2]) - tex3d<R>(G, 0, dn[0], dn[1], dn[2])) /
                            ((R)2 * dX);
                }
                R vis = tex3d<R>(G, 2, uvw[0], uvw[1], uvw[2]);
                for (int c = 0; c < 3; ++c) F[c] -= coeff * gn[c] + vi[c] * vis;
            } else {
                R cgn = std::sqrt(cg[0] * cg[0] + cg[1] * cg[1] + cg[2] * cg[2]);
                if (cgn > (R)P.norm_thresh) {
                    R fc = (R)P.tension_coefficient * cl / cgn;
                    for (int c = 0; c < 3; ++c) F[c] -= fc * cg[c];
                }
            }
            R* ai = &a[3 * (size_t)id];
            ai[0] = F[0] / rhoi;
            ai[1] = F[1] / rhoi - (R)P.gravity;
            ai[2] = F[2] / rhoi;
        }
    }

    // Collision response driven by the distance grid (see file header).
    void collide(R* xp, R* vp, R dt) const {
        R uvw[3];
        tex_coords(G, xp, uvw);
        R D = tex3d<R>(G, 0, uvw[0], uvw[1], uvw[2]);
        if (!(D > (R)0)) return;
        R g[3];
        for (int c = 0; c < 3; ++c) {
            R dX = (R)G.spacing / (R)G.len[c];
            R up[3] = {uvw[0], uvw[1], uvw[2]}, dn[3] = {uvw[0], uvw[1], uvw[2]};
            up[c] += dX;
            dn[c] -= dX;
            g[c] = (tex3d<R>(G, 0, up[0], up[1], up[2]) - tex3d<R>(G, 0, dn[0], dn[1], dn[2])) /
                   ((R)2 * (R)G.spacing);
        }
        R gn = std::sqrt(g[0] * g[0] + g[1] * g[1] + g[2] * g[2]);
        if (!(gn > (R)0)) return;
        R n[3] = {g[0] / gn, g[1] / gn, g[2] / gn};
        R velNorm = std::sqrt(vp[0] * vp[0] + vp[1] * vp[1] + vp[2] * vp[2]);
        R dp = n[0] * vp[0] + n[1] * vp[1] + n[2] * vp[2];
        R e = (R)P.restitution;
        R coeff = (R)1;
        if (e != (R)0 && velNorm > (R)0) coeff += e * D / (dt * velNorm);
        coeff *= dp;
        for (int c = 0; c < 3; ++c) {
            vp[c] -= coeff * n[c];
            xp[c] -= D * n[c];
        }
    }

    // ---- integrate_euler (cu:632-665) [+ collision] ----
    void integrate() {
        const R dt = (R)P.time_step;
#pragma omp parallel for schedule(static) num_threads(threads)
        for (long long i = 0; i < (long long)N; ++i) {
            R* xi = &x[3 * (size_t)i];
            R* vi = &v[3 * (size_t)i];
            const R* ai = &a[3 * (size_t)i];
            for (int c = 0; c < 3; ++c) vi[c] += dt * ai[c];
            for (int c = 0; c < 3; ++c) xi[c] += dt * vi[c];
            if (P.collision) collide(xi, vi, dt);
        }
    }

    // ---- two-scale pass ----

    // shift_state, find_split_particles, find_boundary_particles
    // (cu:730-740, 932-1017)
    void particle_state() {
        for (uint32_t i = 0; i < N; ++i) {
            uint8_t s = (uint8_t)(state[i] << 2);
            const R* xi = &x[3 * (size_t)i];
            bool in = true;
            for (int c = 0; c < 3; ++c)
                in = in && (xi[c] >= (R)P.split_lo[c] && xi[c] <= (R)P.split_hi[c]);
            if (in) s |= 1;
            state[i] = s;
        }
        const R h = (R)P.compact_support;
        std::vector<uint8_t> mark(N, 0);
#pragma omp parallel for schedule(static) num_threads(threads)
        for (long long i = 0; i < (long long)N; ++i) {
            const R* xi = &x[3 * (size_t)i];
            bool hit = false;
            for_neighbours(xi, h, [&](int j) {
                if (hit || !(state[j] & 1)) return;
                const R* xj = &x[3 * (size_t)j];
                R r[3] = {xi[0] - xj[0], xi[1] - xj[1], xi[2] - xj[2]};
                R rn = r[0] * r[0] + r[1] * r[1] + r[2] * r[2];
                if (rn <= h * h) hit = true;
            });
            mark[i] = hit ? 1 : 0;
        }
        for (uint32_t i = 0; i < N; ++i)
            if (mark[i]) state[i] |= 2;
    }

    // check_split_boundary_default + exclusive scans + collect_ids
    // (cu:1092-1162, 2509-2567)
    void collect() {
        split_ids.clear();
        boundary_ids.clear();
        default_ids.clear();
        for (uint32_t i = 0; i < N; ++i) {
            int s = state[i] & 3;
            if (s == 0) default_ids.push_back((int32_t)i);
            else if (s == 2) boundary_ids.push_back((int32_t)i);
            else split_ids.push_back((int32_t)i);
        }
        nsplit = (uint32_t)split_ids.size();
        nboundary = (uint32_t)boundary_ids.size();
        ndefault = (uint32_t)default_ids.size();
        nsub = 8 * (nsplit + nboundary);
        sub_ids.resize(nsub);
        for (uint32_t q = 0; q < nsplit; ++q)
            for (int c = 0; c < 8; ++c) sub_ids[8 * q + c] = 8 * split_ids[q] + c;
        for (uint32_t q = 0; q < nboundary; ++q)
            for (int c = 0; c < 8; ++c) sub_ids[8 * (nsplit + q) + c] = 8 * boundary_ids[q] + c;
    }

    // initialize_sub_particles (cu:1023-1090), every step for every split
    // and boundary parent.  Child c has direction signs
    // (bit2 ? - : +, bit1 ? - : +, bit0 ? - : +) (table at cu:1029-1038).
    void init_subs() {
        sx.assign(3 * (size_t)nsub, 0);
        sv.assign(3 * (size_t)nsub, 0);
        sa.assign(3 * (size_t)nsub, 0);
        sub_cgn_out.assign(nsub, 1e300);
        srho.assign(nsub, 0);
        sp.assign(nsub, 0);
        for (uint32_t q = 0; q < nsub / 8; ++q) {
            int par = sub_ids[8 * q] / 8;
            R radius;
            if (P.sub_placement == 0) {
                R radicand = (R)3 * (R)P.particle_mass / ((R)4 * (R)kPi * rho[par]);
                radius = std::pow(radicand, (R)1 / (R)3);
            } else {
                radius = (R)P.compact_support / (R)4;
            }
            const R s3 = (R)0.577350269;  // SQRT3INV (cu:1025)
            for (int c = 0; c < 8; ++c) {
                R dir[3] = {(R)((c & 4) ? -1 : 1), (R)((c & 2) ? -1 : 1), (R)((c & 1) ? -1 : 1)};
                size_t k = 8 * (size_t)q + c;
                for (int d = 0; d < 3; ++d) {
                    R off = P.sub_placement == 0 ? dir[d] * s3 * radius : dir[d] * radius;
                    sx[3 * k + d] = x[3 * (size_t)par + d] + off;
                    sv[3 * k + d] = v[3 * (size_t)par + d];
                }
            }
        }
    }

    // project_quantities (cu:570-622, 1328-1381): boundary sub-particles take
    // the unweighted mean density and pressure of the coarse particles within
    // h_sub (coarse grid, search radius h_sub).
    void project() {
        const R hs = (R)P.compact_support_sub;
#pragma omp parallel for schedule(static) num_threads(threads)
        for (long long k = 8 * (long long)nsplit; k < (long long)nsub; ++k) {
            const R* xs = &sx[3 * (size_t)k];
            R dsum = 0, psum = 0, cnt = 0;
            for_neighbours(xs, hs, [&](int j) {
                const R* xj = &x[3 * (size_t)j];
                R dx = xs[0] - xj[0], dy = xs[1] - xj[1], dz = xs[2] - xj[2];
                R sq = dx * dx + dy * dy + dz * dz;
                if (sq <= hs * hs) {
                    dsum += rho[j];
                    psum += p[j];
                    cnt += 1;
                }
            });
            if (cnt > 0) {
                srho[k] = dsum / cnt;
                sp[k] = psum / cnt;
            } else {
                srho[k] = (R)P.rest_density;
                sp[k] = 0;
            }
        }
    }

    // compute_sub_particle_hash + copy_array + sort_by_key +
    // compute_cell_start_end on the sub grid (cu:121-153, 2327-2389).
    void sub_grid() {
        skeys_sub.assign(nsub, 0);
        uint32_t f = 0;
        for (uint32_t k = 0; k < nsub; ++k) skeys_sub[k] = sub_cell_key<R>(P, &sx[3 * (size_t)k], f);
        flags |= f;
        sorted_sub.assign(nsub, 0);
        sub_skeys.assign(nsub, 0);
        sub_start.assign(nsubcells, -1);
        sub_end.assign(nsubcells, -1);
        sort_and_cells(nsub, skeys_sub.data(), nsubcells, sub_skeys.data(), sorted_sub.data(),
                       sub_start.data(), sub_end.data());
    }

    template <class F>
    void for_sub_neighbours(const R* xi, R hsearch, F&& fn) const {
        Range r = neighbour_range<R>(P.grid_origin, P.grid_spacing_sub, P.grid_dim_sub, xi, hsearch);
        for (int k = r.lo[2]; k <= r.hi[2]; ++k)
            for (int j = r.lo[1]; j <= r.hi[1]; ++j)
                for (int i = r.lo[0]; i <= r.hi[0]; ++i) {
                    int hash = P.grid_dim_sub[0] * (P.grid_dim_sub[1] * k + j) + i;
                    int s = sub_start[hash], e = sub_end[hash];
                    for (int q = s; q < e; ++q) fn(sorted_sub[q]);
                }
    }

    // compute_sub_particle_density_pressure (cu:259-313, 1423-1457), split
    // sub-particles only, no wall term.
    void sub_density() {
        const R h = (R)P.compact_support_sub;
        const R poly6 = (R)P.poly6_sub;
#pragma omp parallel for schedule(static) num_threads(threads)
        for (long long k = 0; k < 8 * (long long)nsplit; ++k) {
            const R* xi = &sx[3 * (size_t)k];
            R dens = 0;
            for_sub_neighbours(xi, h, [&](int j) {
                const R* xj = &sx[3 * (size_t)j];
                R dx = xj[0] - xi[0], dy = xj[1] - xi[1], dz = xj[2] - xi[2];
                R r = std::sqrt(dx * dx + dy * dy + dz * dz);
                if (r <= h) {
                    R d = h * h - r * r;
                    dens += poly6 * d * d * d;
                }
            });
            dens *= (R)P.sub_particle_mass;
            srho[k] = dens;
            sp[k] = (R)P.gas_stiffness * (dens - (R)P.rest_density);
        }
    }

    // compute_sub_particle_acceleration (cu:477-568, 1558-1643), split subs.
    void sub_acceleration() {
        const R h = (R)P.compact_support_sub;
        const R grad = (R)P.grad_spiky_sub, lapl = (R)P.lapl_visc_sub;
        const R grad2 = (R)P.grad_poly6_sub, lapl2 = (R)P.lapl_poly6_sub;
        const R m = (R)P.sub_particle_mass, mu = (R)P.dynamic_viscosity;
#pragma omp parallel for schedule(static) num_threads(threads)
        for (long long k = 0; k < 8 * (long long)nsplit; ++k) {
            const R* xi = &sx[3 * (size_t)k];
            const R* vi = &sv[3 * (size_t)k];
            R rhoi = srho[k], pi = sp[k], rhoi2 = rhoi * rhoi;
            R F[3] = {0, 0, 0}, cg[3] = {0, 0, 0}, cl = 0;
            for_sub_neighbours(xi, h, [&](int j) {
                const R* xj = &sx[3 * (size_t)j];
                const R* vj = &sv[3 * (size_t)j];
                R rhoj = srho[j], pj = sp[j];
                R r[3] = {xi[0] - xj[0], xi[1] - xj[1], xi[2] - xj[2]};
                R rn = std::sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
                if (rn <= h && rn > (R)0) {
                    R pressure = rhoi * (pi / rhoi2 + pj / (rhoj * rhoj)) * m;
                    R d = (h - rn) * (h - rn);
                    for (int c = 0; c < 3; ++c) F[c] -= pressure * grad * d / rn * r[c];
                    d = h - rn;
                    for (int c = 0; c < 3; ++c) F[c] += mu * (vj[c] - vi[c]) * m / rhoj * lapl * d;
                    d = (h * h - rn * rn) * (h * h - rn * rn);
                    for (int c = 0; c < 3; ++c) cg[c] += m / rhoj * grad2 * d * r[c];
                    d = (h * h - rn * rn) * ((R)3 * h * h - (R)7 * rn * rn);
                    cl += m / rhoj * lapl2 * d;
                }
            });
            R cgn = std::sqrt(cg[0] * cg[0] + cg[1] * cg[1] + cg[2] * cg[2]);
            sub_cgn_out[k] = (double)cgn;
            if (cgn > (R)P.norm_thresh) {
                R fc = (R)P.tension_coefficient * cl / cgn;
                for (int c = 0; c < 3; ++c) F[c] -= fc * cg[c];
            }
            R* ai = &sa[3 * (size_t)k];
            ai[0] = F[0] / rhoi;
            ai[1] = F[1] / rhoi - (R)P.gravity;
            ai[2] = F[2] / rhoi;
        }
    }

    // integrate_sub_particles_euler (cu:667-696) [+ collision].
    void integrate_subs() {
        const R dt = (R)P.time_step_sub;
        for (uint32_t k = 0; k < nsub; ++k) {
            R* xi = &sx[3 * (size_t)k];
            R* vi = &sv[3 * (size_t)k];
            const R* ai = &sa[3 * (size_t)k];
            for (int c = 0; c < 3; ++c) vi[c] += dt * ai[c];
            for (int c = 0; c < 3; ++c) xi[c] += dt * vi[c];
            if (P.collision) collide(xi, vi, dt);
        }
    }

    // AdvanceSubParticles (cu:2119-2143): the current sub-particles alone,
    // no reseeding; projection reads the coarse positions as they are (after
    // the last integration) with the last sort's cell lists and densities.
    void step_sub() {
        project();
        sub_grid();
        sub_density();
        sub_acceleration();
        integrate_subs();
    }

    void step(bool two_scale) {
        hash_sort_cells();
        density();
        acceleration();
        if (two_scale) {
            particle_state();
            collect();
            init_subs();
            project();
            sub_grid();
            sub_density();
            sub_acceleration();
        }
        integrate();
        if (two_scale) integrate_subs();
        for (uint32_t i = 0; i < 3 * N; ++i)
            if (!std::isfinite((double)x[i]) || !std::isfinite((double)v[i])) {
                flags |= ORC_FLAG_NONFINITE;
                break;
            }
        ++steps;
    }

    long read(int field, void* dst, size_t bytes) const {
        auto put_real = [&](const std::vector<R>& src, size_t n) -> long {
            if (bytes < n * sizeof(double)) return -1;
            double* o = (double*)dst;
            for (size_t i = 0; i < n; ++i) o[i] = (double)src[i];
            return (long)n;
        };
        auto put_dbl = [&](const std::vector<double>& src, size_t n) -> long {
            if (bytes < n * sizeof(double)) return -1;
            if (n) std::memcpy(dst, src.data(), n * sizeof(double));
            return (long)n;
        };
        auto put_int = [&](const int32_t* src, size_t n) -> long {
            if (bytes < n * sizeof(int32_t)) return -1;
            if (n) std::memcpy(dst, src, n * sizeof(int32_t));
            return (long)n;
        };
        switch (field) {
        case SPH_FIELD_POSITION: return put_real(x, 3 * (size_t)N);
        case SPH_FIELD_VELOCITY: return put_real(v, 3 * (size_t)N);
        case SPH_FIELD_DENSITY: return put_real(rho, N);
        case SPH_FIELD_PRESSURE: return put_real(p, N);
        case SPH_FIELD_ACCEL: return put_real(a, 3 * (size_t)N);
        case SPH_FIELD_STATE:
            if (bytes < N) return -1;
            if (N) std::memcpy(dst, state.data(), N);
            return (long)N;
        case SPH_FIELD_SORTED_KEYS: return put_int(skeys.data(), N);
        case SPH_FIELD_SORTED_IDS: return put_int(sids.data(), N);
        case SPH_FIELD_CELL_START: return put_int(cstart.data(), ncells);
        case SPH_FIELD_CELL_END: return put_int(cend.data(), ncells);
        case SPH_FIELD_SUB_POSITION: return put_real(sx, 3 * (size_t)nsub);
        case SPH_FIELD_SUB_VELOCITY: return put_real(sv, 3 * (size_t)nsub);
        case SPH_FIELD_SUB_DENSITY: return put_real(srho, nsub);
        case SPH_FIELD_SUB_PRESSURE: return put_real(sp, nsub);
        case SPH_FIELD_SUB_ACCEL: return put_real(sa, 3 * (size_t)nsub);
        case SPH_FIELD_SUB_IDS: return put_int(sub_ids.data(), nsub);
        case SPH_FIELD_SPLIT_IDS: return put_int(split_ids.data(), nsplit);
        case SPH_FIELD_BOUNDARY_IDS: return put_int(boundary_ids.data(), nboundary);
        case SPH_FIELD_DEFAULT_IDS: return put_int(default_ids.data(), ndefault);
        case SPH_FIELD_SUB_SORTED_KEYS: return put_int(sub_skeys.data(), nsub);
        case SPH_FIELD_SUB_CELL_START: return put_int(sub_start.data(), sub_start.size());
        case SPH_FIELD_SUB_CELL_END: return put_int(sub_end.data(), sub_end.size());
        case ORC_FIELD_COLOR_GRAD_NORM: return put_dbl(cgn_out, N);
        case ORC_FIELD_WALL_DISTANCE: return put_dbl(dwall_out, N);
        case ORC_FIELD_SUB_COLOR_GRAD_NORM: return put_dbl(sub_cgn_out, nsub);
        case SPH_FIELD_SUB_SORTED_IDS: {
            if (bytes < nsub * sizeof(int32_t)) return -1;
            int32_t* o = (int32_t*)dst;
            for (uint32_t q = 0; q < nsub; ++q) o[q] = sub_ids[sorted_sub[q]];
            return (long)nsub;
        }
        default: return -1;
        }
    }
};

// ---------------------------------------------------------------------------
// FP64 restatement of the grid builder (SphInComplexShapes.cpp:203-313 for
// the box; same ghost construction for obstacle solids).

double box_sdf(const double p[3], const double c[3], const double e[3]) {
    double t[3] = {p[0] - c[0], p[1] - c[1], p[2] - c[2]};
    double a = std::fabs(t[0]) - e[0], b = std::fabs(t[1]) - e[1], cc = std::fabs(t[2]) - e[2];
    double d = (std::fabs(a) < std::fabs(b)) ? a : b;
    if (std::fabs(d) > std::fabs(cc)) d = cc;
    if (a > 0 || b > 0 || cc > 0) {
        double sq = 0;
        if (a > 0) sq += a * a;
        if (b > 0) sq += b * b;
        if (cc > 0) sq += cc * cc;
        d = std::sqrt(sq);
    }
    return d;
}

double solid_sdf(const sph_solid& s, const double p[3]) {
    if (s.kind == SPH_SOLID_SPHERE) {
        double dx = p[0] - s.center[0], dy = p[1] - s.center[1], dz = p[2] - s.center[2];
        return (double)s.a[0] - std::sqrt(dx * dx + dy * dy + dz * dz);
    }
    if (s.kind == SPH_SOLID_CONE) {
        double R = s.a[0], H = s.a[1];
        double dx = p[0] - s.center[0], dz = p[2] - s.center[2];
        double q = std::sqrt(dx * dx + dz * dz), y = p[1] - s.center[1];
        double bx = std::min(std::max(q, 0.0), R);
        double db = std::hypot(q - bx, y);
        double ex = -R, ey = H;
        double t = std::min(std::max(((q - R) * ex + y * ey) / (ex * ex + ey * ey), 0.0), 1.0);
        double ds = std::hypot(q - (R + t * ex), y - t * ey);
        double dist = std::min(db, ds);
        bool inside = y >= 0 && y <= H && q <= R * (1 - y / H);
        return inside ? dist : -dist;
    }
    if (s.kind == SPH_SOLID_BOX) {
        double c[3] = {s.center[0], s.center[1], s.center[2]}, e[3] = {s.a[0], s.a[1], s.a[2]};
        return -box_sdf(p, c, e);
    }
    return -1e300;
}

double scene_sdf(const sph_scene_desc& d, const double p[3]) {
    double c[3], e[3];
    for (int a = 0; a < 3; ++a) {
        c[a] = 0.5 * ((double)d.box_lo[a] + (double)d.box_hi[a]);
        e[a] = 0.5 * ((double)d.box_hi[a] - (double)d.box_lo[a]);
    }
    double v = box_sdf(p, c, e);
    for (int k = 0; k < d.num_obstacles && k < SPH_MAX_OBSTACLES; ++k) v = std::max(v, solid_sdf(d.obstacles[k], p));
    return v;
}

// FP32 restatements used only for the builder's decisions (which node is
// "taken", which lattice point is a ghost), as the reference decides them in
// float arithmetic.
float box_sdf_f32(const float p[3], const float c[3], const float e[3]) {
    float a = std::fabs(p[0] - c[0]) - e[0], b = std::fabs(p[1] - c[1]) - e[1], cc = std::fabs(p[2] - c[2]) - e[2];
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

float solid_sdf_f32(const sph_solid& s, const float p[3]) {
    if (s.kind == SPH_SOLID_SPHERE) {
        float dx = p[0] - s.center[0], dy = p[1] - s.center[1], dz = p[2] - s.center[2];
        return s.a[0] - std::sqrt(dx * dx + dy * dy + dz * dz);
    }
    if (s.kind == SPH_SOLID_CONE) {
        float R = s.a[0], H = s.a[1];
        float dx = p[0] - s.center[0], dz = p[2] - s.center[2];
        float q = std::sqrt(dx * dx + dz * dz), y = p[1] - s.center[1];
        float bx = std::min(std::max(q, 0.0f), R);
        float db = std::sqrt((q - bx) * (q - bx) + y * y);
        float ex = -R, ey = H;
        float t = std::min(std::max(((q - R) * ex + y * ey) / (ex * ex + ey * ey), 0.0f), 1.0f);
        float sx = R + t * ex, sy = t * ey;
        float ds = std::sqrt((q - sx) * (q - sx) + (y - sy) * (y - sy));
        float dist = std::min(db, ds);
        bool inside = y >= 0.0f && y <= H && q <= R * (1.0f - y / H);
        return inside ? dist : -dist;
    }
    if (s.kind == SPH_SOLID_BOX) return -box_sdf_f32(p, s.center, s.a);
    return -1e30f;
}

float scene_sdf_f32(const sph_scene_desc& d, const float p[3]) {
    float c[3], e[3];
    for (int a = 0; a < 3; ++a) {
        c[a] = 0.5f * (d.box_lo[a] + d.box_hi[a]);
        e[a] = 0.5f * (d.box_hi[a] - d.box_lo[a]);
    }
    float v = box_sdf_f32(p, c, e);
    for (int k = 0; k < d.num_obstacles && k < SPH_MAX_OBSTACLES; ++k) v = std::max(v, solid_sdf_f32(d.obstacles[k], p));
    return v;
}

}  // namespace

struct orc_sim {
    int kind;
    std::unique_ptr<Sim<double>> d;
    std::unique_ptr<Sim<float>> f;
};

extern "C" {

orc_sim* orc_create_f64(const sph_params* prm, const sph_boundary* b, uint32_t n, const double* pos,
                        const double* vel, int real_kind, int threads) {
    if (!prm || !b || (n && !pos)) return nullptr;
    orc_sim* s = new orc_sim();
    s->kind = real_kind;
    if (real_kind == 0) {
        s->d.reset(new Sim<double>());
        s->d->init(*prm, *b, n, pos, vel, threads);
    } else {
        std::vector<float> pf(pos, pos + 3 * (size_t)n), vf;
        if (vel) vf.assign(vel, vel + 3 * (size_t)n);
        s->f.reset(new Sim<float>());
        s->f->init(*prm, *b, n, pf.data(), vel ? vf.data() : nullptr, threads);
    }
    return s;
}

orc_sim* orc_create(const sph_params* prm, const sph_boundary* b, uint32_t n, const float* pos,
                    const float* vel, int real_kind, int threads) {
    if (!prm || !b || (n && !pos)) return nullptr;
    std::vector<double> pd(pos, pos + 3 * (size_t)n), vd;
    if (vel) vd.assign(vel, vel + 3 * (size_t)n);
    return orc_create_f64(prm, b, n, pd.data(), vel ? vd.data() : nullptr, real_kind, threads);
}

void orc_destroy(orc_sim* s) { delete s; }

int orc_step(orc_sim* s, int n_steps, int two_scale) {
    if (!s) return -1;
    auto t0 = std::chrono::steady_clock::now();
    for (int i = 0; i < n_steps; ++i) {
        if (s->d) s->d->step(two_scale != 0);
        else s->f->step(two_scale != 0);
    }
    double sec = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    if (s->d) s->d->last_seconds = sec;
    else s->f->last_seconds = sec;
    return 0;
}

int orc_step_sub(orc_sim* s, int n_steps) {
    if (!s) return -1;
    for (int i = 0; i < n_steps; ++i) {
        if (s->d) s->d->step_sub();
        else s->f->step_sub();
    }
    return 0;
}

long orc_read(orc_sim* s, int field, void* dst, size_t bytes) {
    if (!s) return -1;
    return s->d ? s->d->read(field, dst, bytes) : s->f->read(field, dst, bytes);
}

void orc_counts(orc_sim* s, sph_counts* out) {
    if (!s || !out) return;
    auto fill = [&](auto& S) {
        out->num_particles = S.N;
        out->num_split = S.nsplit;
        out->num_boundary = S.nboundary;
        out->num_default = S.ndefault;
        out->num_sub = S.nsub;
        out->num_steps = S.steps;
    };
    if (s->d) fill(*s->d);
    else fill(*s->f);
}

uint32_t orc_flags(orc_sim* s) { return s ? (s->d ? s->d->flags : s->f->flags) : 0u; }

double orc_last_step_seconds(orc_sim* s) { return s ? (s->d ? s->d->last_seconds : s->f->last_seconds) : 0.0; }

uint32_t orc_keys(const sph_params* p, uint32_t n, const float* pos, int32_t* keys, int real_kind) {
    uint32_t flags = 0;
    for (uint32_t i = 0; i < n; ++i) {
        if (real_kind == 1) {
            keys[i] = cell_key<float>(*p, pos + 3 * (size_t)i, flags);
        } else {
            double x[3] = {pos[3 * (size_t)i], pos[3 * (size_t)i + 1], pos[3 * (size_t)i + 2]};
            keys[i] = cell_key<double>(*p, x, flags);
        }
    }
    return flags;
}

void orc_sort_cells(const sph_params* p, uint32_t n, const int32_t* keys, int32_t* sorted_keys,
                    int32_t* sorted_ids, int32_t* cell_start, int32_t* cell_end) {
    size_t nc = (size_t)p->grid_dim[0] * p->grid_dim[1] * p->grid_dim[2];
    sort_and_cells(n, keys, nc, sorted_keys, sorted_ids, cell_start, cell_end);
}

void orc_tex(const sph_boundary* b, int which, uint32_t n, const double* xyz, double* out) {
    Grid g;
    g.init(*b);
    for (uint32_t i = 0; i < n; ++i) {
        double uvw[3];
        tex_coords(g, xyz + 3 * (size_t)i, uvw);
        out[i] = tex3d<double>(g, which, uvw[0], uvw[1], uvw[2]);
    }
}

void orc_density_brute(const sph_params* P, const sph_boundary* b, uint32_t n, const double* pos, double* rho) {
    Grid g;
    g.init(*b);
    double h = P->compact_support;
    for (uint32_t i = 0; i < n; ++i) {
        double s = 0;
        for (uint32_t j = 0; j < n; ++j) {
            double dx = pos[3 * j] - pos[3 * i], dy = pos[3 * j + 1] - pos[3 * i + 1],
                   dz = pos[3 * j + 2] - pos[3 * i + 2];
            double r2 = dx * dx + dy * dy + dz * dz;
            if (r2 <= h * h) {
                double d = h * h - r2;
                s += (double)P->poly6 * d * d * d;
            }
        }
        double uvw[3];
        tex_coords(g, pos + 3 * (size_t)i, uvw);
        rho[i] = s * P->particle_mass + tex3d<double>(g, 1, uvw[0], uvw[1], uvw[2]);
    }
}

void orc_accel_brute(const sph_params* P, const sph_boundary* b, uint32_t n, const double* pos,
                     const double* vel, const double* rho, double* acc) {
    // Pair sums over all particles; the wall and tension terms come from the
    // same formulas as the cell-grid path, so only the neighbour search
    // differs.
    Sim<double> S;
    S.init(*P, *b, n, pos, vel, 1);
    for (uint32_t i = 0; i < n; ++i) {
        S.rho[i] = rho[i];
        S.p[i] = (double)P->gas_stiffness * (rho[i] - (double)P->rest_density);
    }
    // Make every particle a neighbour of every other: one cell holding all.
    S.P.grid_dim[0] = S.P.grid_dim[1] = S.P.grid_dim[2] = 1;
    S.ncells = 1;
    S.cstart.assign(1, 0);
    S.cend.assign(1, (int32_t)n);
    S.sids.resize(n);
    for (uint32_t i = 0; i < n; ++i) S.sids[i] = (int32_t)i;
    S.acceleration();
    for (size_t i = 0; i < 3 * (size_t)n; ++i) acc[i] = S.a[i];
}

int orc_sub_forces(const sph_params* P, const sph_boundary* b, uint32_t nsub, uint32_t nsplit_subs,
                   const double* pos, const double* vel, const double* rho_in, int threads, double* rho_out,
                   double* acc, double* cgn) {
    // The sub-particle physics of one step on a given sub-particle state:
    // sub grid (compute_sub_particle_hash + sort + compute_cell_start_end,
    // cu:121-153, 2327-2389), density of the split children (cu:259-313)
    // and their acceleration (cu:477-568); the boundary children
    // [nsplit_subs, nsub) carry the densities rho_in (their projection).
    if (!P || !b || nsub % 8 || nsplit_subs % 8 || nsplit_subs > nsub) return -1;
    Sim<double> S;
    S.init(*P, *b, 0, nullptr, nullptr, threads);
    S.nsub = nsub;
    S.nsplit = nsplit_subs / 8;
    S.nboundary = (nsub - nsplit_subs) / 8;
    S.sx.assign(pos, pos + 3 * (size_t)nsub);
    S.sv.assign(vel, vel + 3 * (size_t)nsub);
    S.sa.assign(3 * (size_t)nsub, 0.0);
    S.srho.assign(nsub, 0.0);
    S.sp.assign(nsub, 0.0);
    S.sub_cgn_out.assign(nsub, 1e300);
    for (uint32_t k = nsplit_subs; k < nsub; ++k) {
        S.srho[k] = rho_in[k];
        S.sp[k] = (double)P->gas_stiffness * (rho_in[k] - (double)P->rest_density);
    }
    S.sub_grid();
    S.sub_density();
    S.sub_acceleration();
    for (uint32_t k = 0; k < nsub; ++k) {
        rho_out[k] = S.srho[k];
        cgn[k] = S.sub_cgn_out[k];
        for (int c = 0; c < 3; ++c) acc[3 * (size_t)k + c] = S.sa[3 * (size_t)k + c];
    }
    return (int)S.flags;
}

int orc_scene_boundary(const sph_scene_desc* d, double* distance, double* density, double* viscosity) {
    // Decisions (lattice extents, node/ghost positions, inclusion tests, the
    // taken-node test |d| < cs, splat index boxes) follow the reference's FP32
    // expressions, as the cell keys do; accumulated values are FP64.
    if (!d) return -1;
    int dims[3] = {d->boundary_dims[0], d->boundary_dims[1], d->boundary_dims[2]};
    const float* o = d->boundary_origin;
    const float dx = d->boundary_spacing, r0 = d->rest_distance, cs = d->boundary_support;
    size_t nn = (size_t)dims[0] * dims[1] * dims[2];
    std::vector<char> taken(nn, 0);
    std::vector<double> D(nn, -(double)r0), W(nn, 0.0), V(nn, 0.0);
    auto idx = [&](int i, int j, int k) { return (size_t)i + (size_t)dims[0] * ((size_t)j + (size_t)dims[1] * k); };
    auto node = [&](int i, int j, int k, float p[3]) {
        p[0] = o[0] + dx * (float)i;
        p[1] = o[1] + dx * (float)j;
        p[2] = o[2] + dx * (float)k;
    };
    for (int k = 0; k < dims[2]; ++k)
        for (int j = 0; j < dims[1]; ++j)
            for (int i = 0; i < dims[0]; ++i) {
                float pf[3];
                node(i, j, k, pf);
                double p[3] = {pf[0], pf[1], pf[2]};
                if (std::fabs(scene_sdf_f32(*d, pf)) < cs) {
                    double v = scene_sdf(*d, p);
                    if (std::fabs(v) > r0) v = v > 0 ? r0 : -r0;
                    D[idx(i, j, k)] = v;
                    taken[idx(i, j, k)] = 1;
                }
            }
    double mass = d->particle_mass > 0 ? (double)d->particle_mass
                                       : (double)d->rest_density * std::pow((double)d->lattice_spacing, 3);
    float massf = d->particle_mass > 0 ? d->particle_mass
                                       : d->rest_density * d->lattice_spacing * d->lattice_spacing * d->lattice_spacing;
    float gsp = d->ghost_spacing > 0 ? d->ghost_spacing : std::pow(massf / d->rest_density, 1.0f / 3.0f);
    double csd = cs;
    double poly = 315.0 / (64.0 * kPi * std::pow(csd, 9));
    double visk = 45.0 / (kPi * std::pow(csd, 6));
    auto coord = [&](float v, int a) {
        int c = (int)((v - o[a]) / dx);
        return std::min(std::max(c, 0), dims[a] - 1);
    };
    auto splat = [&](const float g[3]) {
        int lo[3], hi[3];
        for (int a = 0; a < 3; ++a) {
            lo[a] = coord(g[a] - cs, a);
            hi[a] = coord(g[a] + cs, a);
        }
        for (int k = lo[2]; k <= hi[2]; ++k)
            for (int j = lo[1]; j <= hi[1]; ++j)
                for (int i = lo[0]; i <= hi[0]; ++i) {
                    size_t q = idx(i, j, k);
                    if (!taken[q]) continue;
                    float pf[3];
                    node(i, j, k, pf);
                    double r2 = 0;
                    for (int a = 0; a < 3; ++a) r2 += ((double)g[a] - pf[a]) * ((double)g[a] - pf[a]);
                    double df = csd * csd - r2;
                    if (df >= 0) W[q] += poly * df * df * df;
                    double r = std::sqrt(r2);
                    if (r <= csd) V[q] += visk * (csd - r);
                }
    };
    float c[3], e[3];
    for (int a = 0; a < 3; ++a) {
        c[a] = 0.5f * (d->box_lo[a] + d->box_hi[a]);
        e[a] = 0.5f * (d->box_hi[a] - d->box_lo[a]);
    }
    auto inside_open = [&](const float p[3]) {
        return !(std::fabs(p[0] - c[0]) >= e[0] || std::fabs(p[1] - c[1]) >= e[1] ||
                 std::fabs(p[2] - c[2]) >= e[2]);
    };
    {
        // SetBox ghost lattice (SphInComplexShapes.cpp:232-260)
        float dd = cs + gsp;
        float oe[3] = {e[0] + dd, e[1] + dd, e[2] + dd};
        float s0[3] = {c[0] - oe[0], c[1] - oe[1], c[2] - oe[2]};
        unsigned mx[3];
        for (int a = 0; a < 3; ++a) mx[a] = (unsigned)(2 * oe[a] / gsp);
        for (unsigned k = 0; k <= mx[2]; ++k)
            for (unsigned j = 0; j <= mx[1]; ++j)
                for (unsigned i = 0; i <= mx[0]; ++i) {
                    float p[3] = {s0[0] + (float)i * gsp, s0[1] + (float)j * gsp, s0[2] + (float)k * gsp};
                    if (!inside_open(p)) splat(p);
                }
    }
    for (int ob = 0; ob < d->num_obstacles && ob < SPH_MAX_OBSTACLES; ++ob) {
        const sph_solid& s = d->obstacles[ob];
        float lo[3], hi[3];
        if (s.kind == SPH_SOLID_SPHERE) {
            for (int a = 0; a < 3; ++a) { lo[a] = s.center[a] - s.a[0]; hi[a] = s.center[a] + s.a[0]; }
        } else if (s.kind == SPH_SOLID_CONE) {
            lo[0] = s.center[0] - s.a[0]; hi[0] = s.center[0] + s.a[0];
            lo[1] = s.center[1];          hi[1] = s.center[1] + s.a[1];
            lo[2] = s.center[2] - s.a[0]; hi[2] = s.center[2] + s.a[0];
        } else if (s.kind == SPH_SOLID_BOX) {
            for (int a = 0; a < 3; ++a) { lo[a] = s.center[a] - s.a[a]; hi[a] = s.center[a] + s.a[a]; }
        } else {
            continue;
        }
        float dd = cs + gsp;
        unsigned n3[3];
        for (int a = 0; a < 3; ++a) {
            lo[a] -= dd;
            hi[a] += dd;
            n3[a] = (unsigned)((hi[a] - lo[a]) / gsp);
        }
        for (unsigned k = 0; k <= n3[2]; ++k)
            for (unsigned j = 0; j <= n3[1]; ++j)
                for (unsigned i = 0; i <= n3[0]; ++i) {
                    float p[3] = {lo[0] + (float)i * gsp, lo[1] + (float)j * gsp, lo[2] + (float)k * gsp};
                    float v = solid_sdf_f32(s, p);
                    if (v >= 0.0f && v <= dd && inside_open(p)) splat(p);
                }
    }
    double vs = (double)d->viscosity * mass / (double)d->rest_density;
    for (size_t q = 0; q < nn; ++q) {
        if (distance) distance[q] = D[q];
        if (density) density[q] = W[q] * mass;
        if (viscosity) viscosity[q] = V[q] * vs;
    }
    return 0;
}

void orc_scene_sdf(const sph_scene_desc* d, uint32_t n, const double* xyz, double* out) {
    for (uint32_t i = 0; i < n; ++i) out[i] = scene_sdf(*d, xyz + 3 * (size_t)i);
}

}  // extern "C"

extern "C" {
// Layout of the ABI structs as g++ sees include/sph.h (checked against the
// ctypes mirror by the tests).
void orc_struct_layout(size_t* out) {
    out[0] = sizeof(sph_params);
    out[1] = sizeof(sph_boundary);
    out[2] = sizeof(sph_scene);
    out[3] = sizeof(sph_scene_desc);
    out[4] = sizeof(sph_counts);
    out[5] = sizeof(sph_solid);
    out[6] = offsetof(sph_params, max_sub_particles);
    out[7] = offsetof(sph_scene, boundary);
    out[8] = offsetof(sph_scene_desc, sub_placement);
    out[9] = offsetof(sph_boundary, viscosity);
    out[10] = sizeof(sph_map_solid);
    out[11] = offsetof(sph_map_solid, distance);
    out[12] = offsetof(sph_map_solid, solid_hi);
}
}
