"""Pins the FP64 mesh signed-distance oracle (oracle/mesh_oracle.cpp) and the
host side of the mesh path (OBJ reader, fit, domain, save/load) on the CPU.

No reference output exists for this path (no saved boundary map ships with
the reference, and its WildMagic5 distance/intersection routines are absent),
so the oracle is pinned by closed forms, brute force and an analytic box."""
import os

import numpy as np
import pytest

from cudasphnew_amd import meshsdf as M
from oracle import oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))
MESHES = np.load(os.path.join(HERE, "golden", "meshes.npz"))


def box_mesh(lo, hi):
    x0, y0, z0 = lo
    x1, y1, z1 = hi
    v = np.array([[x0, y0, z0], [x1, y0, z0], [x1, y1, z0], [x0, y1, z0],
                  [x0, y0, z1], [x1, y0, z1], [x1, y1, z1], [x0, y1, z1]], np.float32)
    f = np.array([[0, 2, 1], [0, 3, 2], [4, 5, 6], [4, 6, 7], [0, 1, 5], [0, 5, 4],
                  [2, 3, 7], [2, 7, 6], [1, 2, 6], [1, 6, 5], [0, 4, 7], [0, 7, 3]], np.uint32)
    return v, f


def test_point_triangle_closed_forms():
    tri = [0, 0, 0, 1, 0, 0, 0, 1, 0]
    pts = np.array([[0.2, 0.2, 0.5], [-1, -1, 0], [0.5, -1, 0.3], [1, 1, 0], [2, 0, 0], [0.1, 0.1, 0]])
    want = [0.5, np.sqrt(2), np.sqrt(1 + 0.09), 1 / np.sqrt(2), 1.0, 0.0]
    assert np.allclose(O.point_triangle(pts, tri), want, atol=1e-15)


def test_point_triangle_brute_force():
    rng = np.random.default_rng(3)
    for _ in range(20):
        tri = rng.normal(size=(3, 3))
        pts = rng.normal(size=(50, 3)) * 2
        d = O.point_triangle(pts, tri)
        # dense barycentric samples: their minimum bounds the distance from
        # above and exceeds it by at most the sampling resolution
        u, v = np.meshgrid(np.linspace(0, 1, 201), np.linspace(0, 1, 201))
        m = u + v <= 1
        s = tri[0] + u[m, None] * (tri[1] - tri[0]) + v[m, None] * (tri[2] - tri[0])
        dmin = np.sqrt(((pts[:, None, :] - s[None]) ** 2).sum(-1)).min(1)
        res = np.linalg.norm(tri[1] - tri[0]) / 200 + np.linalg.norm(tri[2] - tri[0]) / 200
        assert np.all(d <= dmin + 1e-12)
        assert np.all(dmin - d <= res)


def test_inside_parity_on_watertight_meshes_including_projected_vertices():
    v, f = MESHES["icosphere_vertices"], MESHES["icosphere_faces"]
    c = v.mean(0)
    r = np.linalg.norm(v - c, axis=1)
    rng = np.random.default_rng(1)
    dirs = rng.normal(size=(400, 3))
    dirs /= np.linalg.norm(dirs, axis=1)[:, None]
    inner = c + 0.7 * r.min() * dirs * rng.random((400, 1))
    outer = c + 1.3 * r.max() * dirs
    assert O.inside(v, f, inner).all()
    assert not O.inside(v, f, outer).any()
    # rays through projected vertices and edges: x, y of the vertices, z at
    # the centre (inside) and above (outside)
    on_v = np.column_stack([v[:, 0], v[:, 1], np.full(len(v), c[2])])
    deep = np.linalg.norm(on_v - c, axis=1) < 0.7 * r.min()
    assert O.inside(v, f, on_v[deep]).all()
    above = np.column_stack([v[:, 0], v[:, 1], np.full(len(v), c[2] + 2 * r.max())])
    assert not O.inside(v, f, above).any()


def test_box_mesh_sdf_matches_analytic_box():
    lo, hi = (0.1, 0.2, 0.3), (0.9, 0.7, 0.6)
    v, f = box_mesh(lo, hi)
    rest, cs, dx = 0.1, 0.05, 0.02
    dims, dlo, dhi = M.domain(v, f, dx)
    d, n, p = O.mesh_sdf(v, f, rest, cs, dx, dims, dlo, dhi)
    nz, ny, nx = d.shape
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    X = np.stack([dlo[0] + x * np.float32(dx), dlo[1] + y * np.float32(dx), dlo[2] + z * np.float32(dx)], -1)
    X = X.astype(np.float32).astype(np.float64)
    c = (np.array(lo) + np.array(hi)) / 2
    e = (np.array(hi) - np.array(lo)) / 2
    q = np.abs(X - c) - e
    sdf = np.linalg.norm(np.maximum(q, 0), axis=-1) + np.minimum(q.max(-1), 0)
    inside = q.max(-1) < 0
    maxd = max(rest, cs)
    assert np.all(p[np.abs(sdf) < maxd - 1e-6] == 1)
    assert np.all(p[np.abs(sdf) > maxd + 1e-6] == 0)
    # inside the box the distance is to the nearest face; the box mesh's
    # triangles reproduce it exactly, with the sign from the parity test
    ok = p == 1
    assert np.allclose(d[ok], sdf[ok], atol=1e-6)
    clear = ok & (np.abs(sdf) > 1e-6)
    assert np.array_equal(d[clear] < 0, inside[clear])
    assert np.all(d[p == 0] == np.float32(rest))
    # normals: unit, pointing from the surface to the node
    nn = np.linalg.norm(n[ok & (np.abs(sdf) > 1e-6)], axis=-1)
    assert np.allclose(nn, 1.0, atol=1e-12)


def test_domain_rule():
    v, f = MESHES["cone_vertices"], MESHES["cone_faces"]
    dx = np.float32(0.017 / 3)
    dims, lo, hi = M.domain(v, f, float(dx))
    v1, v2 = v.min(0), v.max(0)
    diff = np.ceil((v2 - v1) * (np.float32(1) / dx)).astype(np.float32)
    assert dims == tuple(int(x) + 1 for x in diff)
    assert np.array_equal(lo, v1)
    assert np.allclose(hi, v1 + diff * dx, rtol=0, atol=1e-6)


def test_fit_maps_bounding_box(tmp_path):
    v = MESHES["bunny_vertices"]
    w = M.fit(v, (-0.75, -0.75, -0.75), (0.75, 0.75, 0.75))
    assert np.allclose(w.min(0), -0.75, atol=1e-6) and np.allclose(w.max(0), 0.75, atol=1e-5)


def test_obj_reader_round_trip(tmp_path):
    v, f = MESHES["icosphere_vertices"], MESHES["icosphere_faces"]
    path = tmp_path / "m.obj"
    with open(path, "w") as fh:
        fh.write("# test\no obj\n")
        for p in v:
            fh.write(f"v {p[0]:.9g} {p[1]:.9g} {p[2]:.9g}\n")
        for t in f[:-2]:
            fh.write(f"f {t[0] + 1} {t[1] + 1} {t[2] + 1}\n")
        # slash form and a quad that fans into the last two triangles' shape
        fh.write(f"f {f[-2][0] + 1}/1/1 {f[-2][1] + 1}/1/1 {f[-2][2] + 1}/1/1\n")
        fh.write(f"f {f[-1][0] + 1} {f[-1][1] + 1} {f[-1][2] + 1} {f[0][0] + 1}\n")
    rv, rf = M.read_obj(str(path))
    assert np.array_equal(rv, v)
    assert np.array_equal(rf[:-1], f)
    assert list(rf[-1]) == [f[-1][0], f[-1][2], f[0][0]]


def test_save_load_round_trip(tmp_path):
    rng = np.random.default_rng(2)
    dims = (7, 5, 4)
    shape = (dims[2], dims[1], dims[0])
    cfg = M.config(0.034, 0.017, 0.017 / 3)
    p = (rng.random(shape) < 0.4).astype(np.uint8)
    d = np.where(p == 1, rng.normal(size=shape), cfg.rest_distance).astype(np.float32)
    n = np.where(p[..., None] == 1, rng.normal(size=shape + (3,)), 0).astype(np.float32)
    bm = {"dims": dims, "lo": np.array([0.1, 0.2, 0.3], np.float32), "hi": np.array([1, 2, 3], np.float32),
          "distance": d, "normal": n, "present": p}
    M.save(str(tmp_path / "b.txt"), cfg, bm)
    cfg2, bm2 = M.load(str(tmp_path / "b.txt"))
    assert bm2["dims"] == dims
    assert np.array_equal(bm2["present"], p)
    assert np.array_equal(bm2["distance"], d)
    assert np.array_equal(bm2["normal"], n)
    assert abs(cfg2.dx - cfg.dx) < 1e-9
