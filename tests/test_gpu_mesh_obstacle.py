"""Obstacles from meshes: a mesh's boundary map, generated on the GPU on a
grid grown past the mesh's box (sph_boundary_map_generate_on), becomes an
obstacle of a scene (sph_scene_boundary_maps) and the step runs with the
resulting grids.

Bars: the generated field on the grown grid against the FP64 oracle as in
test_gpu_mesh.py; the grids of a finely subdivided icosphere against those
of the analytic sphere it approximates (polyhedral error <= 4e-4 for the
radius below, trilinear error ~1e-4); particles of a dam break that runs
into the mesh sphere stay out of it as far as they stay out of the analytic
one, also with the learned boundary trained on the scene with the map."""
import ctypes as C
import os

import numpy as np
import pytest

from cudasphnew_amd import Field
from cudasphnew_amd import meshsdf as M
from cudasphnew_amd import neural as NN
from cudasphnew_amd import scenes as S
from cudasphnew_amd._native import SphSolid
from cudasphnew_amd.simulation import ParticleSimulation
from oracle import oracle as O

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
MESHES = np.load(os.path.join(HERE, "golden", "meshes.npz"))

# beside the dam block's +z face (the block spans 0.038 .. 0.348 in x and z)
CENTER, RADIUS = (0.19, 0.10, 0.475), 0.1037


def unit_icosphere(levels):
    """The fixture icosphere (its vertices lie on the unit sphere about the
    centre of their box), each triangle split in four `levels` times with
    the new vertices pushed onto the sphere."""
    v = MESHES["icosphere_vertices"].astype(np.float64)
    v = v - 0.5 * (v.min(0) + v.max(0))
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    f = MESHES["icosphere_faces"].astype(np.int64)
    for _ in range(levels):
        verts = list(v)
        mid = {}

        def midpoint(a, b):
            key = (min(a, b), max(a, b))
            if key not in mid:
                m = verts[a] + verts[b]
                verts.append(m / np.linalg.norm(m))
                mid[key] = len(verts) - 1
            return mid[key]

        nf = []
        for a, b, c in f:
            ab, bc, ca = midpoint(a, b), midpoint(b, c), midpoint(c, a)
            nf += [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
        v, f = np.array(verts), np.array(nf, np.int64)
    return v, f.astype(np.uint32)


def sphere_mesh(levels=3):
    v, f = unit_icosphere(levels)
    return (np.array(CENTER) + RADIUS * v).astype(np.float32), f


def analytic(desc):
    d = S.SceneDesc.from_buffer_copy(desc)
    d.num_obstacles = 1
    d.obstacles[0] = SphSolid(1, (C.c_float * 3)(*CENTER), (C.c_float * 3)(RADIUS, 0, 0))
    return d


def mesh_obstacle(desc):
    v, f = sphere_mesh()
    bm = S.mesh_obstacle(v, f, desc)
    # the ghost lattice of the sphere starts from the sphere's box; the
    # inscribed mesh's own box is a hair smaller
    bm["mesh_lo"] = np.array(CENTER, np.float32) - np.float32(RADIUS)
    bm["mesh_hi"] = np.array(CENTER, np.float32) + np.float32(RADIUS)
    return bm, v, f


def test_generated_on_a_grown_grid_matches_oracle():
    v = M.fit(MESHES["cone_vertices"], (-0.5, -0.5, -0.5), (0.5, 0.5, 0.5))
    f = MESHES["cone_faces"]
    rest, cs, dx = 0.06, 0.03, 0.01
    dims, lo = M.grown_domain(v, dx, 0.07)
    bm = M.generate_on(v, f, M.config(rest, cs, dx), dims, lo)
    assert bm["dims"] == dims and np.all(bm["lo"] < v.min(0) - 0.069)
    d, n, p = O.mesh_sdf(v, f, rest, cs, dx, bm["dims"], bm["lo"], bm["hi"])
    maxd = float(np.float32(max(rest, cs)))
    edge = np.abs(np.abs(d) - maxd) < 1e-5
    assert np.array_equal(bm["present"][~edge], p[~edge])
    both = (bm["present"] == 1) & (p == 1)
    assert np.abs(bm["distance"][both] - d[both]).max() <= 1e-5
    clear = both & (np.abs(d) > 1e-5)
    assert np.array_equal(bm["distance"][clear] < 0, d[clear] < 0)
    # the band is whole on every side: present nodes reach past the mesh's
    # box by maxDist (to the node spacing) along each axis
    zz, yy, xx = np.nonzero(bm["present"])
    for a, idx in enumerate((xx, yy, zz)):
        coord = bm["lo"][a] + dx * idx
        assert coord.min() < v[:, a].min() - maxd + 1.01 * dx
        assert coord.max() > v[:, a].max() + maxd - 1.01 * dx
    # a domain that misses part of the mesh is generated without error (the
    # triangles outside it are skipped) and agrees with the oracle where it
    # lies
    lo2 = (v.min(0) + 0.4).astype(np.float32)
    bm2 = M.generate_on(v, f, M.config(rest, cs, dx), (20, 30, 25), lo2)
    d2, _, p2 = O.mesh_sdf(v, f, rest, cs, dx, bm2["dims"], bm2["lo"], bm2["hi"])
    edge2 = np.abs(np.abs(d2) - maxd) < 1e-5
    assert np.array_equal(bm2["present"][~edge2], p2[~edge2])


def test_mesh_sphere_grids_match_the_analytic_sphere():
    plain = S.config_desc("dam16k")
    bm, v, f = mesh_obstacle(plain)
    # polyhedral error of the mesh: distance of the face centres to the sphere
    fc = v[f].astype(np.float64).mean(1)
    sag = RADIUS - np.linalg.norm(fc - np.array(CENTER), axis=1).min()
    assert sag < 4e-4
    A = S.boundary_of(analytic(plain))
    B = S.boundary_of(plain, [bm])
    E = S.boundary_of(plain)
    diff = np.abs(A[0] - B[0])
    moved = diff > sag + 2e-4
    print(f"mesh sphere: {f.shape[0]} faces, polyhedral error {sag:.2e}, map {bm['dims']} in {bm['ms']:.2f} ms; "
          f"distance max {diff[~moved].max():.2e} off {moved.sum()} nodes")
    assert moved.sum() <= 40
    for g in (A[0], B[0]):
        assert np.all(np.abs(np.abs(g[moved]) - plain.boundary_support) < 2e-3)
    for a, b, e in zip(A[1:], B[1:], E[1:]):
        share = np.abs(a - e)
        rel = np.abs(a - b).sum() / share.sum()
        print(f"  density/viscosity: error {rel:.4f} of the sphere's share")
        assert rel <= 0.05


def run_against_sphere(scene, net=None, val=None):
    """6,000 steps of 1e-4 s (the dam's front runs over the sphere): closest
    approach of a particle to the sphere's centre, the largest number of
    particles within 1.5 h of its surface at a check, the attach report and
    the boundary sources of the last step."""
    h = float(scene.desc.h)
    with ParticleSimulation(scene).Init() as sim:
        rep = sim.attach_neural(net, val=val) if net is not None else None
        closest, touching = np.inf, 0
        for _ in range(20):
            sim.Advance(300)
            pos = sim.positions()
            assert np.all(np.isfinite(pos))
            r = np.linalg.norm(pos - np.array(CENTER, np.float32), axis=1)
            closest = min(closest, float(r.min()))
            touching = max(touching, int((r < RADIUS + 1.5 * h).sum()))
        src = np.bincount(sim.read(Field.WALL_SOURCE), minlength=3)
    return closest, touching, rep, src


def test_dam_break_around_a_mesh_obstacle():
    plain = S.config_desc("dam16k")
    bm, _, _ = mesh_obstacle(plain)
    runs = {}
    for name, desc, maps in (("analytic", analytic(plain), None), ("mesh", plain, [bm])):
        scene = S.scene_from_desc(desc, maps)
        r0 = np.linalg.norm(scene.positions - np.array(CENTER, np.float32), axis=1)
        assert r0.min() > RADIUS + 0.02
        runs[name] = run_against_sphere(scene)[:2]
        print(f"{name}: closest approach to the centre {runs[name][0]:.4f} (radius {RADIUS}), "
              f"up to {runs[name][1]} particles within 1.5 h of the surface")
    # the fluid runs against the sphere in both runs (the boundary force
    # holds particles about r0 = h off a wall, as off the container's walls)
    for name in runs:
        assert runs[name][1] > 50
        assert runs[name][0] < RADIUS + 1.25 * plain.h
    # and keeps off the mesh sphere as it keeps off the analytic one
    assert abs(runs["mesh"][0] - runs["analytic"][0]) < 2e-3
    assert runs["mesh"][0] > RADIUS


def test_learned_boundary_around_a_mesh_obstacle():
    """The learned boundary of a scene whose obstacle comes from a mesh:
    trained on the scene's grid samples with the map (the recipe of the
    plain scenes), validated within 0.05 h on its band, and the dam break
    run with it keeps off the sphere as the grid run does."""
    plain = S.config_desc("dam16k")
    bm, _, _ = mesh_obstacle(plain)
    net = NN.train_for_scene(plain, maps=[bm])
    val = NN.validation_samples(plain, maps=[bm])
    scene = S.scene_from_desc(plain, [bm])
    grid = run_against_sphere(scene)
    learned = run_against_sphere(scene, net, val)
    rep, src = learned[2], learned[3]
    print(f"grid: closest {grid[0]:.4f}, touching {grid[1]}; learned: closest {learned[0]:.4f}, "
          f"touching {learned[1]}; report {rep}; sources {src}")
    assert rep["mean_error"] < 0.05 * plain.h
    # distance and penetration-weighted normal bars: a mesh's facets leave
    # more bricks on the grid than an analytic scene's
    assert rep["bricks_fallback"] <= 0.15 * rep["bricks_checked"]
    assert rep["mean_normal_error"] < 0.05
    assert src[1] > 0.2 * scene.n          # the network is the wall source of most wall particles
    assert learned[1] > 50 and learned[0] > RADIUS
    # the learned normals and distances differ from the grid's inside the
    # penalty zone: the fluid keeps the same distance to 0.15 h (measured
    # 0.07 h)
    assert abs(learned[0] - grid[0]) < 0.15 * plain.h
