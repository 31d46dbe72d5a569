// mesh.hip -- signed distance field of a triangle mesh on a grid, the
// BoundaryMap of the reference (boundary_map.cpp:60-100 AddCanvas,
// 369-490 computeSignedDistanceField / computeGridCoordinates,
// point_in_mesh_test.cpp PointInMeshTest) rebuilt for the GPU.
//
// * Domain: the mesh's bounding box v1..v2, dims = ceil((v2 - v1)/dx) + 1 per
//   axis, v2 moved to v1 + (dims - 1) dx (AddCanvas).
// * Unsigned distance: every triangle visits the nodes of its bounding box
//   grown by maxDist + dx and clamped to the domain (computeGridCoordinates);
//   a node keeps the smallest distance <= maxDist, ties to the lowest triangle
//   index (the reference keeps the first triangle, `prev > dist`).  On the
//   GPU the (triangle, node-chunk) pairs are independent work items and the
//   minimum is an atomicMin over (distance bits << 32 | triangle).
// * Normal: (x - closest point) / |x - closest point| of the winning
//   triangle, not flipped by the sign (compute_distance_point_triangle); 0
//   when the node lies on the triangle.
// * Sign: negative inside.  The reference casts a ray of random direction
//   and retries when it grazes an edge; here the ray is +z and grazing is
//   resolved by a fixed rule: the node's (x, y) is tested against each
//   triangle's projection with orientation signs computed exactly in double
//   and the top-left rule for points on an edge, so a closed mesh's
//   crossings are counted once each and the parity is deterministic.
//   Triangles are found through a uniform xy grid of their projected bounds.
//
// Closest point on a triangle: the Voronoi-region method of Ericson,
// "Real-Time Collision Detection" (2005), section 5.1.5.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "sph.h"

namespace sph {
namespace mesh {

struct Tri {
    float3 a, b, c;
};

__host__ __device__ inline float3 sub3(float3 u, float3 v) { return make_float3(u.x - v.x, u.y - v.y, u.z - v.z); }
__host__ __device__ inline float dot3(float3 u, float3 v) { return u.x * v.x + u.y * v.y + u.z * v.z; }

// Closest point of triangle (a, b, c) to p (Ericson 5.1.5).
__device__ inline float3 closest_point(float3 p, float3 a, float3 b, float3 c) {
    const float3 ab = sub3(b, a), ac = sub3(c, a), ap = sub3(p, a);
    const float d1 = dot3(ab, ap), d2 = dot3(ac, ap);
    if (d1 <= 0.0f && d2 <= 0.0f) return a;
    const float3 bp = sub3(p, b);
    const float d3 = dot3(ab, bp), d4 = dot3(ac, bp);
    if (d3 >= 0.0f && d4 <= d3) return b;
    const float vc = d1 * d4 - d3 * d2;
    if (vc <= 0.0f && d1 >= 0.0f && d3 <= 0.0f) {
        const float v = d1 / (d1 - d3);
        return make_float3(a.x + v * ab.x, a.y + v * ab.y, a.z + v * ab.z);
    }
    const float3 cp = sub3(p, c);
    const float d5 = dot3(ab, cp), d6 = dot3(ac, cp);
    if (d6 >= 0.0f && d5 <= d6) return c;
    const float vb = d5 * d2 - d1 * d6;
    if (vb <= 0.0f && d2 >= 0.0f && d6 <= 0.0f) {
        const float w = d2 / (d2 - d6);
        return make_float3(a.x + w * ac.x, a.y + w * ac.y, a.z + w * ac.z);
    }
    const float va = d3 * d6 - d5 * d4;
    if (va <= 0.0f && (d4 - d3) >= 0.0f && (d5 - d6) >= 0.0f) {
        const float w = (d4 - d3) / ((d4 - d3) + (d5 - d6));
        return make_float3(b.x + w * (c.x - b.x), b.y + w * (c.y - b.y), b.z + w * (c.z - b.z));
    }
    const float denom = 1.0f / (va + vb + vc);
    const float v = vb * denom, w = vc * denom;
    return make_float3(a.x + ab.x * v + ac.x * w, a.y + ab.y * v + ac.y * w, a.z + ab.z * v + ac.z * w);
}

struct Domain {
    float lo[3];
    float dx;
    int dims[3];
    float maxdist;
};

// Node box of a triangle (computeGridCoordinates, boundary_map.cpp:498-536):
// bounds grown by maxDist + dx, clamped to the domain, truncated.  False
// when the grown bounds miss the domain (possible on a caller-given grid),
// so that no node lies within maxDist of the triangle.
inline bool triangle_box(const Domain& D, const float hi[3], const float* t1, const float* t2, const float* t3,
                         unsigned cmin[3], unsigned cmax[3]) {
    const float delta = D.maxdist + D.dx;
    for (int a = 0; a < 3; ++a) {
        float mn = std::min(t1[a], std::min(t2[a], t3[a])) - delta;
        float mx = std::max(t1[a], std::max(t2[a], t3[a])) + delta;
        if (mx < D.lo[a] || mn > hi[a]) return false;
        mn = std::max(mn, D.lo[a]);
        mx = std::min(mx, hi[a]);
        cmin[a] = std::min((unsigned)((mn - D.lo[a]) / D.dx), (unsigned)D.dims[a] - 1u);
        cmax[a] = std::min((unsigned)((mx - D.lo[a]) / D.dx), (unsigned)D.dims[a] - 1u);
    }
    return true;
}

// One work item: a run of up to kChunk nodes of one triangle's node box.
struct Item {
    uint32_t tri;
    uint32_t first;         // first node of the run in the box's i-fastest order
    uint32_t box0[3];       // box origin (node coordinates)
    uint32_t boxn[3];       // box extent (nodes)
};
constexpr int kChunk = 1024;

__global__ __launch_bounds__(256) void k_udf(const Item* __restrict__ items, const Tri* __restrict__ tris, Domain D,
                                             unsigned long long* __restrict__ best) {
    const Item it = items[blockIdx.x];
    const Tri t = tris[it.tri];
    const uint32_t total = it.boxn[0] * it.boxn[1] * it.boxn[2];
    for (int q = threadIdx.x; q < kChunk; q += blockDim.x) {
        const uint32_t r = it.first + (uint32_t)q;
        if (r >= total) break;
        const uint32_t i = it.box0[0] + r % it.boxn[0];
        const uint32_t j = it.box0[1] + (r / it.boxn[0]) % it.boxn[1];
        const uint32_t k = it.box0[2] + r / (it.boxn[0] * it.boxn[1]);
        const float3 x = make_float3(D.lo[0] + i * D.dx, D.lo[1] + j * D.dx, D.lo[2] + k * D.dx);
        const float3 c = closest_point(x, t.a, t.b, t.c);
        const float3 d = sub3(x, c);
        const float dist = sqrtf(dot3(d, d));
        if (dist <= D.maxdist) {
            const unsigned long long key = ((unsigned long long)__float_as_uint(dist) << 32) | it.tri;
            atomicMin(&best[i + (size_t)D.dims[0] * (j + (size_t)D.dims[1] * k)], key);
        }
    }
}

// Exact-sign 2D orientation of (p, a, b) from float inputs: differences of
// floats are exact in double and their products fit in 53 bits.
__device__ inline double orient2(double px, double py, float ax, float ay, float bx, float by) {
    return ((double)ax - px) * ((double)by - py) - ((double)ay - py) * ((double)bx - px);
}

// Top-left rule for a point on edge (a -> b) of a counter-clockwise triangle:
// the edge owns its points if it is a "left" edge (going down) or a "top"
// edge (horizontal, going left).
__device__ inline bool owns_edge(float ax, float ay, float bx, float by) {
    return (by < ay) || (by == ay && bx < ax);
}

// +z ray from p crosses triangle t (projection test with the top-left rule,
// crossing strictly above p).
__device__ inline bool crosses(const Tri& t, float3 p) {
    const double px = p.x, py = p.y;
    float ax = t.a.x, ay = t.a.y, bx = t.b.x, by = t.b.y, cx = t.c.x, cy = t.c.y;
    const double area = orient2((double)ax, (double)ay, bx, by, cx, cy);
    if (area == 0.0) return false;                     // edge-on in projection
    float3 A = t.a, B = t.b, C = t.c;
    if (area < 0.0) {                                   // make it counter-clockwise
        float tx = bx, ty = by;
        bx = cx; by = cy; cx = tx; cy = ty;
        float3 T = B; B = C; C = T;
    }
    const double w0 = orient2(px, py, bx, by, cx, cy);  // edge b->c
    const double w1 = orient2(px, py, cx, cy, ax, ay);  // edge c->a
    const double w2 = orient2(px, py, ax, ay, bx, by);  // edge a->b
    if (w0 < 0.0 || w1 < 0.0 || w2 < 0.0) return false;
    if (w0 == 0.0 && !owns_edge(bx, by, cx, cy)) return false;
    if (w1 == 0.0 && !owns_edge(cx, cy, ax, ay)) return false;
    if (w2 == 0.0 && !owns_edge(ax, ay, bx, by)) return false;
    const double s = w0 + w1 + w2;
    const double z = (w0 * A.z + w1 * B.z + w2 * C.z) / s;   // plane height at (px, py)
    return z > (double)p.z;
}

// Distance, normal and sign of every node; absent nodes (farther than
// maxDist from every triangle) get the reference map's default content
// (restDistance, 0, 0, 0) (BoundaryMap::Load, boundary_map.cpp:243-246).
__global__ void k_finish(const unsigned long long* __restrict__ best, const Tri* __restrict__ tris, Domain D,
                         const uint32_t* __restrict__ cell_start, const uint32_t* __restrict__ cell_tris,
                         int hashn, float hx0, float hy0, float hdx, float hdy, float bbz0, float bbz1,
                         float rest_distance, float* __restrict__ dist, float* __restrict__ normal,
                         uint8_t* __restrict__ present) {
    const size_t n = (size_t)D.dims[0] * D.dims[1] * D.dims[2];
    const size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    const unsigned long long b = best[q];
    if (b == ~0ull) {
        dist[q] = rest_distance;
        normal[3 * q] = normal[3 * q + 1] = normal[3 * q + 2] = 0.0f;
        present[q] = 0;
        return;
    }
    const uint32_t i = (uint32_t)(q % D.dims[0]), j = (uint32_t)((q / D.dims[0]) % D.dims[1]),
                   k = (uint32_t)(q / ((size_t)D.dims[0] * D.dims[1]));
    const float3 x = make_float3(D.lo[0] + i * D.dx, D.lo[1] + j * D.dx, D.lo[2] + k * D.dx);
    const Tri t = tris[(uint32_t)(b & 0xFFFFFFFFull)];
    const float3 c = closest_point(x, t.a, t.b, t.c);
    const float3 d = sub3(x, c);
    const float len = sqrtf(dot3(d, d));
    float d_out = __uint_as_float((uint32_t)(b >> 32));
    float3 nrm = len > 0.0f ? make_float3(d.x / len, d.y / len, d.z / len) : make_float3(0.0f, 0.0f, 0.0f);
    // sign: parity of the +z ray's crossings (points outside the mesh's
    // bounding box are outside)
    bool inside = false;
    const int hi_ = (int)floorf((x.x - hx0) / hdx), hj = (int)floorf((x.y - hy0) / hdy);
    if (hi_ >= 0 && hi_ < hashn && hj >= 0 && hj < hashn && x.z >= bbz0 && x.z <= bbz1) {
        const uint32_t cell = (uint32_t)(hi_ + hashn * hj);
        uint32_t cnt = 0;
        for (uint32_t e = cell_start[cell]; e < cell_start[cell + 1]; ++e) cnt += crosses(tris[cell_tris[e]], x);
        inside = cnt & 1u;
    }
    dist[q] = inside ? -d_out : d_out;
    normal[3 * q] = nrm.x;
    normal[3 * q + 1] = nrm.y;
    normal[3 * q + 2] = nrm.z;
    present[q] = 1;
}

}  // namespace mesh
}  // namespace sph

using namespace sph::mesh;

static bool mesh_ok(const sph_mesh* m) {
    if (!m || !m->vertices || !m->faces || m->num_vertices == 0 || m->num_faces == 0) return false;
    for (size_t f = 0; f < 3 * (size_t)m->num_faces; ++f)
        if (m->faces[f] >= m->num_vertices) return false;
    return true;
}

static void bounding_box(const sph_mesh* m, float lo[3], float hi[3]) {
    for (int a = 0; a < 3; ++a) lo[a] = hi[a] = m->vertices[a];
    for (uint32_t v = 0; v < m->num_vertices; ++v)
        for (int a = 0; a < 3; ++a) {
            lo[a] = std::min(lo[a], m->vertices[3 * v + a]);
            hi[a] = std::max(hi[a], m->vertices[3 * v + a]);
        }
}

extern "C" {

sph_status sph_obj_read(const char* path, float* vertices, uint32_t* num_vertices, uint32_t* faces,
                        uint32_t* num_faces) {
    if (!path || !num_vertices || !num_faces) return SPH_INVALID_ARG;
    std::ifstream f(path);
    if (!f) return SPH_INVALID_ARG;
    const uint32_t cap_v = *num_vertices, cap_f = *num_faces;
    uint32_t nv = 0, nf = 0;
    std::string line;
    while (std::getline(f, line)) {
        if (line.size() < 2) continue;
        if (line[0] == 'v' && (line[1] == ' ' || line[1] == '\t')) {
            std::istringstream s(line.substr(2));
            float x, y, z;
            if (!(s >> x >> y >> z)) return SPH_INVALID_ARG;
            if (vertices && nv < cap_v) {
                vertices[3 * nv] = x;
                vertices[3 * nv + 1] = y;
                vertices[3 * nv + 2] = z;
            }
            ++nv;
        } else if (line[0] == 'f' && (line[1] == ' ' || line[1] == '\t')) {
            // "f a b c" or "f a/t/n ...", 1-based; polygons are fanned
            std::istringstream s(line.substr(2));
            std::vector<long> idx;
            std::string tok;
            while (s >> tok) {
                long v = std::strtol(tok.c_str(), nullptr, 10);
                if (v < 0) v = (long)nv + 1 + v;     // relative index
                if (v < 1) return SPH_INVALID_ARG;
                idx.push_back(v - 1);
            }
            if (idx.size() < 3) return SPH_INVALID_ARG;
            for (size_t t = 1; t + 1 < idx.size(); ++t) {
                if (faces && nf < cap_f) {
                    faces[3 * nf] = (uint32_t)idx[0];
                    faces[3 * nf + 1] = (uint32_t)idx[t];
                    faces[3 * nf + 2] = (uint32_t)idx[t + 1];
                }
                ++nf;
            }
        }
    }
    *num_vertices = nv;
    *num_faces = nf;
    if ((vertices && nv > cap_v) || (faces && nf > cap_f)) return SPH_CAPACITY_EXCEEDED;
    return SPH_OK;
}

sph_status sph_mesh_fit(float* vertices, uint32_t n, const float lo[3], const float hi[3]) {
    if (!vertices || n == 0 || !lo || !hi) return SPH_INVALID_ARG;
    sph_mesh m = {vertices, n, nullptr, 0};
    float blo[3], bhi[3];
    bounding_box(&m, blo, bhi);
    // TriangleMesh::fit (triangle_mesh.cpp): scale each axis about the box
    // centre to the target extent, then move the box's corner to lo
    float ratio[3], mid[3];
    for (int a = 0; a < 3; ++a) {
        float ext = bhi[a] - blo[a];
        if (!(ext > 0.0f)) return SPH_INVALID_ARG;
        ratio[a] = (hi[a] - lo[a]) / ext;
        mid[a] = blo[a] + 0.5f * ext;
    }
    for (uint32_t v = 0; v < n; ++v)
        for (int a = 0; a < 3; ++a) vertices[3 * v + a] = (vertices[3 * v + a] - mid[a]) * ratio[a] + mid[a];
    bounding_box(&m, blo, bhi);
    for (uint32_t v = 0; v < n; ++v)
        for (int a = 0; a < 3; ++a) vertices[3 * v + a] += lo[a] - blo[a];
    return SPH_OK;
}

sph_status sph_boundary_map_domain(const sph_mesh* m, float dx, int32_t dims[3], float lo[3], float hi[3]) {
    if (!mesh_ok(m) || !(dx > 0.0f) || !dims || !lo || !hi) return SPH_INVALID_ARG;
    float v1[3], v2[3];
    bounding_box(m, v1, v2);
    for (int a = 0; a < 3; ++a) {
        float diff = std::ceil((v2[a] - v1[a]) * (1.0f / dx));
        dims[a] = (int32_t)diff + 1;
        lo[a] = v1[a];
        hi[a] = v1[a] + diff * dx;
    }
    return SPH_OK;
}

static bool config_ok(const sph_boundary_map_config* cfg) {
    return cfg && cfg->dx > 0.0f && cfg->compact_support > 0.0f && cfg->rest_distance >= 0.0f &&
           std::isfinite(cfg->dx) && std::isfinite(cfg->compact_support) && std::isfinite(cfg->rest_distance);
}

// Both generators: the field over domain D (D.lo .. hi, D.dims nodes).
static sph_status generate_on(const sph_mesh* m, const sph_boundary_map_config* cfg, Domain D, const float hi[3],
                              float* distance, float* normal, uint8_t* present, float* ms) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) {
        (void)hipGetLastError();
        return SPH_NO_DEVICE;
    }
    D.dx = cfg->dx;
    D.maxdist = std::max(cfg->rest_distance, cfg->compact_support);
    const size_t nodes = (size_t)D.dims[0] * D.dims[1] * D.dims[2];
    if (nodes > (size_t)1 << 31) return SPH_INVALID_ARG;
    // triangles and work items
    std::vector<Tri> tris(m->num_faces);
    std::vector<Item> items;
    for (uint32_t f = 0; f < m->num_faces; ++f) {
        const float* p1 = m->vertices + 3 * m->faces[3 * f];
        const float* p2 = m->vertices + 3 * m->faces[3 * f + 1];
        const float* p3 = m->vertices + 3 * m->faces[3 * f + 2];
        tris[f].a = make_float3(p1[0], p1[1], p1[2]);
        tris[f].b = make_float3(p2[0], p2[1], p2[2]);
        tris[f].c = make_float3(p3[0], p3[1], p3[2]);
        unsigned cmin[3], cmax[3];
        if (!triangle_box(D, hi, p1, p2, p3, cmin, cmax)) continue;
        Item it;
        it.tri = f;
        uint32_t total = 1;
        for (int a = 0; a < 3; ++a) {
            it.box0[a] = cmin[a];
            it.boxn[a] = cmax[a] >= cmin[a] ? cmax[a] - cmin[a] + 1 : 0;
            total *= it.boxn[a];
        }
        for (uint32_t first = 0; first < total; first += kChunk) {
            it.first = first;
            items.push_back(it);
        }
    }
    // xy hash of the triangles' projected bounds over the mesh's box
    float blo[3], bhi[3];
    bounding_box(m, blo, bhi);
    const int hn = std::max(1, std::min(256, (int)std::sqrt((double)m->num_faces)));
    const float hdx = std::max((bhi[0] - blo[0]) / hn, 1e-12f), hdy = std::max((bhi[1] - blo[1]) / hn, 1e-12f);
    std::vector<uint32_t> cnt((size_t)hn * hn + 1, 0), start((size_t)hn * hn + 1, 0), list;
    auto cell_range = [&](const Tri& t, int& i0, int& i1, int& j0, int& j1) {
        float x0 = std::min(t.a.x, std::min(t.b.x, t.c.x)), x1 = std::max(t.a.x, std::max(t.b.x, t.c.x));
        float y0 = std::min(t.a.y, std::min(t.b.y, t.c.y)), y1 = std::max(t.a.y, std::max(t.b.y, t.c.y));
        i0 = std::max(0, (int)std::floor((x0 - blo[0]) / hdx));
        i1 = std::min(hn - 1, (int)std::floor((x1 - blo[0]) / hdx));
        j0 = std::max(0, (int)std::floor((y0 - blo[1]) / hdy));
        j1 = std::min(hn - 1, (int)std::floor((y1 - blo[1]) / hdy));
    };
    for (const Tri& t : tris) {
        int i0, i1, j0, j1;
        cell_range(t, i0, i1, j0, j1);
        for (int j = j0; j <= j1; ++j)
            for (int i = i0; i <= i1; ++i) ++cnt[i + (size_t)hn * j];
    }
    for (size_t c = 0; c < (size_t)hn * hn; ++c) start[c + 1] = start[c] + cnt[c];
    list.resize(start[(size_t)hn * hn]);
    std::vector<uint32_t> fill(start.begin(), start.end() - 1);
    for (uint32_t f = 0; f < m->num_faces; ++f) {
        int i0, i1, j0, j1;
        cell_range(tris[f], i0, i1, j0, j1);
        for (int j = j0; j <= j1; ++j)
            for (int i = i0; i <= i1; ++i) list[fill[i + (size_t)hn * j]++] = f;
    }
    // device
    Tri* d_tris = nullptr;
    Item* d_items = nullptr;
    unsigned long long* d_best = nullptr;
    uint32_t *d_start = nullptr, *d_list = nullptr;
    float *d_dist = nullptr, *d_norm = nullptr;
    uint8_t* d_present = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    bool ok = hipMalloc(&d_tris, sizeof(Tri) * tris.size()) == hipSuccess &&
              hipMalloc(&d_items, sizeof(Item) * std::max<size_t>(items.size(), 1)) == hipSuccess &&
              hipMalloc(&d_best, sizeof(unsigned long long) * nodes) == hipSuccess &&
              hipMalloc(&d_start, sizeof(uint32_t) * start.size()) == hipSuccess &&
              hipMalloc(&d_list, sizeof(uint32_t) * std::max<size_t>(list.size(), 1)) == hipSuccess &&
              hipMalloc(&d_dist, sizeof(float) * nodes) == hipSuccess &&
              hipMalloc(&d_norm, sizeof(float) * 3 * nodes) == hipSuccess &&
              hipMalloc(&d_present, nodes) == hipSuccess && hipEventCreate(&e0) == hipSuccess &&
              hipEventCreate(&e1) == hipSuccess;
    if (ok) {
        ok = hipMemcpy(d_tris, tris.data(), sizeof(Tri) * tris.size(), hipMemcpyHostToDevice) == hipSuccess &&
             (items.empty() ||
              hipMemcpy(d_items, items.data(), sizeof(Item) * items.size(), hipMemcpyHostToDevice) == hipSuccess) &&
             hipMemcpy(d_start, start.data(), sizeof(uint32_t) * start.size(), hipMemcpyHostToDevice) == hipSuccess &&
             (list.empty() ||
              hipMemcpy(d_list, list.data(), sizeof(uint32_t) * list.size(), hipMemcpyHostToDevice) == hipSuccess) &&
             hipMemset(d_best, 0xFF, sizeof(unsigned long long) * nodes) == hipSuccess;
    }
    if (ok) {
        (void)hipEventRecord(e0, nullptr);
        if (!items.empty())
            hipLaunchKernelGGL(k_udf, dim3((unsigned)items.size()), dim3(256), 0, nullptr, d_items, d_tris, D, d_best);
        hipLaunchKernelGGL(k_finish, dim3((unsigned)((nodes + 255) / 256)), dim3(256), 0, nullptr, d_best, d_tris, D,
                           d_start, d_list, hn, blo[0], blo[1], hdx, hdy, blo[2], bhi[2], cfg->rest_distance, d_dist,
                           d_norm, d_present);
        (void)hipEventRecord(e1, nullptr);
        ok = hipEventSynchronize(e1) == hipSuccess && hipGetLastError() == hipSuccess;
        if (ok && ms) (void)hipEventElapsedTime(ms, e0, e1);
        ok = ok && hipMemcpy(distance, d_dist, sizeof(float) * nodes, hipMemcpyDeviceToHost) == hipSuccess &&
             hipMemcpy(normal, d_norm, sizeof(float) * 3 * nodes, hipMemcpyDeviceToHost) == hipSuccess &&
             hipMemcpy(present, d_present, nodes, hipMemcpyDeviceToHost) == hipSuccess;
    }
    void* ptrs[] = {d_tris, d_items, d_best, d_start, d_list, d_dist, d_norm, d_present};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (!ok) {
        (void)hipGetLastError();
        return SPH_DEVICE_ERROR;
    }
    return SPH_OK;
}

sph_status sph_boundary_map_generate(const sph_mesh* m, const sph_boundary_map_config* cfg, float* distance,
                                     float* normal, uint8_t* present, float* ms) {
    if (!mesh_ok(m) || !config_ok(cfg) || !distance || !normal || !present) return SPH_INVALID_ARG;
    Domain D;
    float hi[3];
    sph_status st = sph_boundary_map_domain(m, cfg->dx, D.dims, D.lo, hi);
    if (st != SPH_OK) return st;
    return generate_on(m, cfg, D, hi, distance, normal, present, ms);
}

sph_status sph_boundary_map_generate_on(const sph_mesh* m, const sph_boundary_map_config* cfg,
                                        const int32_t dims[3], const float lo[3], float* distance, float* normal,
                                        uint8_t* present, float* ms) {
    if (!mesh_ok(m) || !config_ok(cfg) || !dims || !lo || !distance || !normal || !present) return SPH_INVALID_ARG;
    Domain D;
    float hi[3];
    for (int a = 0; a < 3; ++a) {
        if (dims[a] < 1 || !std::isfinite(lo[a])) return SPH_INVALID_ARG;
        D.dims[a] = dims[a];
        D.lo[a] = lo[a];
        hi[a] = lo[a] + (float)(dims[a] - 1) * cfg->dx;
    }
    return generate_on(m, cfg, D, hi, distance, normal, present, ms);
}

// BoundaryMap::Save text format (boundary_map.cpp:161-221): restDistance,
// compactSupport, maxDist, dx, domain v1, v2, iMax, jMax, kMax, total
// samples, number of nodes, then per node i, j, k, distance, normal xyz.
// Nodes are written in ascending linear index (the reference writes them in
// insertion order; its Load reads coordinates, not order).
sph_status sph_boundary_map_save(const char* path, const sph_boundary_map_config* cfg, const int32_t dims[3],
                                 const float lo[3], const float hi[3], const float* distance, const float* normal,
                                 const uint8_t* present) {
    if (!path || !cfg || !dims || !lo || !hi || !distance || !normal || !present) return SPH_INVALID_ARG;
    FILE* f = std::fopen(path, "w");
    if (!f) return SPH_INVALID_ARG;
    const size_t n = (size_t)dims[0] * dims[1] * dims[2];
    size_t cnt = 0;
    for (size_t q = 0; q < n; ++q) cnt += present[q] != 0;
    std::fprintf(f, "%.9g\n%.9g\n%.9g\n%.9g\n", cfg->rest_distance, cfg->compact_support,
                 std::max(cfg->rest_distance, cfg->compact_support), cfg->dx);
    for (int a = 0; a < 3; ++a) std::fprintf(f, "%.9g\n", lo[a]);
    for (int a = 0; a < 3; ++a) std::fprintf(f, "%.9g\n", hi[a]);
    std::fprintf(f, "%d\n%d\n%d\n%zu\n%zu\n", dims[0], dims[1], dims[2], n, cnt);
    for (size_t q = 0; q < n; ++q) {
        if (!present[q]) continue;
        size_t i = q % dims[0], j = (q / dims[0]) % dims[1], k = q / ((size_t)dims[0] * dims[1]);
        std::fprintf(f, "%zu\n%zu\n%zu\n%.9g\n%.9g\n%.9g\n%.9g\n", i, j, k, distance[q], normal[3 * q],
                     normal[3 * q + 1], normal[3 * q + 2]);
    }
    bool ok = std::fclose(f) == 0;
    return ok ? SPH_OK : SPH_DEVICE_ERROR;
}

// BoundaryMap::Load (boundary_map.cpp:223-279): header, then nodes by
// coordinate; absent nodes hold (restDistance, 0, 0, 0).  Call with NULL
// arrays to read dims/lo/hi/config only.
sph_status sph_boundary_map_load(const char* path, sph_boundary_map_config* cfg, int32_t dims[3], float lo[3],
                                 float hi[3], float* distance, float* normal, uint8_t* present) {
    if (!path || !cfg || !dims || !lo || !hi) return SPH_INVALID_ARG;
    std::ifstream f(path);
    if (!f) return SPH_INVALID_ARG;
    float maxdist;
    size_t total, cnt;
    if (!(f >> cfg->rest_distance >> cfg->compact_support >> maxdist >> cfg->dx)) return SPH_INVALID_ARG;
    for (int a = 0; a < 3; ++a) f >> lo[a];
    for (int a = 0; a < 3; ++a) f >> hi[a];
    if (!(f >> dims[0] >> dims[1] >> dims[2] >> total >> cnt)) return SPH_INVALID_ARG;
    const size_t n = (size_t)dims[0] * dims[1] * dims[2];
    if (total != n) return SPH_INVALID_ARG;
    if (!distance && !normal && !present) return SPH_OK;
    if (!distance || !normal || !present) return SPH_INVALID_ARG;
    for (size_t q = 0; q < n; ++q) {
        distance[q] = cfg->rest_distance;
        normal[3 * q] = normal[3 * q + 1] = normal[3 * q + 2] = 0.0f;
        present[q] = 0;
    }
    for (size_t e = 0; e < cnt; ++e) {
        size_t i, j, k;
        float d, nx, ny, nz;
        if (!(f >> i >> j >> k >> d >> nx >> ny >> nz)) return SPH_INVALID_ARG;
        if (i >= (size_t)dims[0] || j >= (size_t)dims[1] || k >= (size_t)dims[2]) return SPH_INVALID_ARG;
        size_t q = i + dims[0] * (j + (size_t)dims[1] * k);
        distance[q] = d;
        normal[3 * q] = nx;
        normal[3 * q + 1] = ny;
        normal[3 * q + 2] = nz;
        present[q] = 1;
    }
    return SPH_OK;
}

}  // extern "C"
