// mesh_oracle.cpp -- TEST INFRASTRUCTURE: FP64 restatement of the mesh
// signed distance field (BoundaryMap::AddCanvas / computeSignedDistanceField,
// boundary_map.cpp:60-100, 369-536; PointInMeshTest, point_in_mesh_test.cpp).
// Only tests/ use it, as the checker; the product never links it.
//
// Node selection per triangle uses the reference's FP32 box arithmetic
// (computeGridCoordinates); distances are FP64 (closest point on a triangle,
// Ericson, "Real-Time Collision Detection" 5.1.5); a node keeps the smallest
// distance <= maxDist, ties to the first triangle (`prev > dist`,
// boundary_map.cpp:427).  Sign: parity of the +z ray's crossings, each
// triangle's projection tested with orientation signs exact in double and
// the top-left rule for points on an edge (the reference's random ray and
// retry on grazing hits have no reproducible value; this rule fixes one).
// Parity unpinned against the reference's WildMagic5 routines
// (DistPoint3Triangle3f, IntrRay3Triangle3f: implementations absent): the
// distance is pinned by closed forms and brute force in tests/.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <vector>

namespace {

struct V3 {
    double x, y, z;
};
V3 sub(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
double dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
V3 madd(V3 a, V3 d, double t) { return {a.x + t * d.x, a.y + t * d.y, a.z + t * d.z}; }

V3 closest(V3 p, V3 a, V3 b, V3 c) {
    V3 ab = sub(b, a), ac = sub(c, a), ap = sub(p, a);
    double d1 = dot(ab, ap), d2 = dot(ac, ap);
    if (d1 <= 0 && d2 <= 0) return a;
    V3 bp = sub(p, b);
    double d3 = dot(ab, bp), d4 = dot(ac, bp);
    if (d3 >= 0 && d4 <= d3) return b;
    double vc = d1 * d4 - d3 * d2;
    if (vc <= 0 && d1 >= 0 && d3 <= 0) return madd(a, ab, d1 / (d1 - d3));
    V3 cp = sub(p, c);
    double d5 = dot(ab, cp), d6 = dot(ac, cp);
    if (d6 >= 0 && d5 <= d6) return c;
    double vb = d5 * d2 - d1 * d6;
    if (vb <= 0 && d2 >= 0 && d6 <= 0) return madd(a, ac, d2 / (d2 - d6));
    double va = d3 * d6 - d5 * d4;
    if (va <= 0 && (d4 - d3) >= 0 && (d5 - d6) >= 0) return madd(b, sub(c, b), (d4 - d3) / ((d4 - d3) + (d5 - d6)));
    double den = 1.0 / (va + vb + vc);
    double v = vb * den, w = vc * den;
    return {a.x + ab.x * v + ac.x * w, a.y + ab.y * v + ac.y * w, a.z + ab.z * v + ac.z * w};
}

double orient2(double px, double py, float ax, float ay, float bx, float by) {
    return ((double)ax - px) * ((double)by - py) - ((double)ay - py) * ((double)bx - px);
}
bool owns(float ax, float ay, float bx, float by) { return (by < ay) || (by == ay && bx < ax); }

bool crosses(const float* A, const float* B, const float* C, float px_, float py_, float pz) {
    const double px = px_, py = py_;
    float ax = A[0], ay = A[1], bx = B[0], by = B[1], cx = C[0], cy = C[1];
    double area = orient2(ax, ay, bx, by, cx, cy);
    if (area == 0.0) return false;
    const float *P = A, *Q = B, *R = C;
    if (area < 0.0) {
        std::swap(bx, cx);
        std::swap(by, cy);
        std::swap(Q, R);
    }
    double w0 = orient2(px, py, bx, by, cx, cy), w1 = orient2(px, py, cx, cy, ax, ay), w2 = orient2(px, py, ax, ay, bx, by);
    if (w0 < 0 || w1 < 0 || w2 < 0) return false;
    if (w0 == 0 && !owns(bx, by, cx, cy)) return false;
    if (w1 == 0 && !owns(cx, cy, ax, ay)) return false;
    if (w2 == 0 && !owns(ax, ay, bx, by)) return false;
    double s = w0 + w1 + w2;
    double z = (w0 * P[2] + w1 * Q[2] + w2 * R[2]) / s;
    return z > (double)pz;
}

}  // namespace

extern "C" {

// dims/lo/hi as the product's sph_boundary_map_domain; outputs over the
// dims grid: distance (signed, rest_distance where absent), normal xyz,
// present.
void orc_mesh_sdf(const float* vert, uint32_t nv, const uint32_t* faces, uint32_t nf, float dx, float rest,
                  float support, const int32_t dims[3], const float lo[3], const float hi[3], double* dist,
                  double* normal, uint8_t* present) {
    (void)nv;
    const float maxd = std::max(rest, support);
    const size_t n = (size_t)dims[0] * dims[1] * dims[2];
    std::vector<double> best(n, INFINITY);
    std::vector<int64_t> who(n, -1);
    for (uint32_t f = 0; f < nf; ++f) {
        const float* t[3] = {vert + 3 * faces[3 * f], vert + 3 * faces[3 * f + 1], vert + 3 * faces[3 * f + 2]};
        unsigned cmin[3], cmax[3];
        const float delta = maxd + dx;
        for (int a = 0; a < 3; ++a) {
            float mn = std::min(t[0][a], std::min(t[1][a], t[2][a])) - delta;
            float mx = std::max(t[0][a], std::max(t[1][a], t[2][a])) + delta;
            mn = std::max(mn, lo[a]);
            mx = std::min(mx, hi[a]);
            cmin[a] = (unsigned)((mn - lo[a]) / dx);
            cmax[a] = std::min((unsigned)((mx - lo[a]) / dx), (unsigned)dims[a] - 1u);
        }
        V3 A{t[0][0], t[0][1], t[0][2]}, B{t[1][0], t[1][1], t[1][2]}, C{t[2][0], t[2][1], t[2][2]};
#pragma omp parallel for collapse(2) schedule(static)
        for (unsigned k = cmin[2]; k <= cmax[2]; ++k)
            for (unsigned j = cmin[1]; j <= cmax[1]; ++j)
                for (unsigned i = cmin[0]; i <= cmax[0]; ++i) {
                    V3 x{(double)(lo[0] + i * dx), (double)(lo[1] + j * dx), (double)(lo[2] + k * dx)};
                    V3 c = closest(x, A, B, C);
                    V3 d = sub(x, c);
                    double dd = std::sqrt(dot(d, d));
                    size_t q = i + (size_t)dims[0] * (j + (size_t)dims[1] * k);
                    if (dd <= maxd && dd < best[q]) {
                        best[q] = dd;
                        who[q] = f;
                    }
                }
    }
    float blo[3], bhi[3];
    for (int a = 0; a < 3; ++a) blo[a] = bhi[a] = vert[a];
    for (uint32_t v = 0; v < nv; ++v)
        for (int a = 0; a < 3; ++a) {
            blo[a] = std::min(blo[a], vert[3 * v + a]);
            bhi[a] = std::max(bhi[a], vert[3 * v + a]);
        }
#pragma omp parallel for schedule(dynamic, 4096)
    for (size_t q = 0; q < n; ++q) {
        if (who[q] < 0) {
            dist[q] = rest;
            normal[3 * q] = normal[3 * q + 1] = normal[3 * q + 2] = 0.0;
            present[q] = 0;
            continue;
        }
        size_t i = q % dims[0], j = (q / dims[0]) % dims[1], k = q / ((size_t)dims[0] * dims[1]);
        const float xf[3] = {lo[0] + i * dx, lo[1] + j * dx, lo[2] + k * dx};
        const uint32_t f = (uint32_t)who[q];
        V3 A{vert[3 * faces[3 * f]], vert[3 * faces[3 * f] + 1], vert[3 * faces[3 * f] + 2]};
        V3 B{vert[3 * faces[3 * f + 1]], vert[3 * faces[3 * f + 1] + 1], vert[3 * faces[3 * f + 1] + 2]};
        V3 C{vert[3 * faces[3 * f + 2]], vert[3 * faces[3 * f + 2] + 1], vert[3 * faces[3 * f + 2] + 2]};
        V3 x{xf[0], xf[1], xf[2]};
        V3 d = sub(x, closest(x, A, B, C));
        double len = std::sqrt(dot(d, d));
        bool inside = false;
        if (xf[0] >= blo[0] && xf[0] <= bhi[0] && xf[1] >= blo[1] && xf[1] <= bhi[1] && xf[2] >= blo[2] &&
            xf[2] <= bhi[2]) {
            unsigned cnt = 0;
            for (uint32_t g = 0; g < nf; ++g)
                cnt += crosses(vert + 3 * faces[3 * g], vert + 3 * faces[3 * g + 1], vert + 3 * faces[3 * g + 2], xf[0],
                               xf[1], xf[2]);
            inside = cnt & 1u;
        }
        dist[q] = inside ? -best[q] : best[q];
        for (int a = 0; a < 3; ++a) normal[3 * q + a] = len > 0 ? (&d.x)[a] / len : 0.0;
        present[q] = 1;
    }
}

// Distance of points to one triangle (closed-form checks).
void orc_point_triangle(const double* p, size_t n, const double* tri, double* out) {
    V3 A{tri[0], tri[1], tri[2]}, B{tri[3], tri[4], tri[5]}, C{tri[6], tri[7], tri[8]};
    for (size_t i = 0; i < n; ++i) {
        V3 x{p[3 * i], p[3 * i + 1], p[3 * i + 2]};
        V3 d = sub(x, closest(x, A, B, C));
        out[i] = std::sqrt(dot(d, d));
    }
}

// Parity of the +z ray from each point over all triangles (1 = inside).
void orc_inside(const float* vert, const uint32_t* faces, uint32_t nf, const float* p, size_t n, uint8_t* out) {
#pragma omp parallel for schedule(static)
    for (size_t i = 0; i < n; ++i) {
        unsigned cnt = 0;
        for (uint32_t g = 0; g < nf; ++g)
            cnt += crosses(vert + 3 * faces[3 * g], vert + 3 * faces[3 * g + 1], vert + 3 * faces[3 * g + 2],
                           p[3 * i], p[3 * i + 1], p[3 * i + 2]);
        out[i] = cnt & 1u;
    }
}

}  // extern "C"
