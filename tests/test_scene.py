"""Scene construction and the boundary grid builder (host code, no GPU).

The builder is pinned by the reference's own golden slices distances.ppm and
viscosities.ppm (decoded into tests/golden/ppm_fixtures.npz), by the FP64
oracle builder, and by closed forms of the signed distance.
"""
import os

import numpy as np
import pytest

import cudasphnew_amd as S
from oracle import oracle as O
from tests.helpers import ppm_slice

HERE = os.path.dirname(os.path.abspath(__file__))
FIX = np.load(os.path.join(HERE, "golden", "ppm_fixtures.npz"))


@pytest.fixture(scope="module")
def ex01():
    return S.example01_boundary()


def test_example01_derivation(ex01):
    # Example01: N = 34^3, h = (3*0.125*20/(4 pi N))^(1/3), grid spacing h/4
    # over (-0.8,-0.6,-0.4)..(0.8,0.6,0.4) (particle_simulation.cu:1745-1777,
    # 2596-2607; SphInComplexShapes.cpp:49-54)
    N = 34 ** 3
    h = (3 * 0.125 * 20 / (4 * np.pi * N)) ** (1 / 3)
    assert ex01["rest_distance"] == pytest.approx(h, rel=1e-6)
    assert ex01["spacing"] == pytest.approx(h / 4, rel=1e-6)
    assert ex01["dims"] == (259, 194, 130)
    assert FIX["distances"].shape == (194, 259, 3)


def test_distances_ppm_exact(ex01):
    nz = ex01["dims"][2]
    D = ex01["distance"][nz // 2]
    r0 = np.float32(ex01["rest_distance"])
    img = ppm_slice(D / r0)
    assert np.array_equal(img, FIX["distances"])


def test_viscosities_ppm(ex01):
    nz = ex01["dims"][2]
    r0 = np.float32(ex01["rest_distance"])
    V = ex01["viscosity"]
    vmax = np.float32(V.max())
    k = nz // 2
    taken = ex01["distance"][k] != -r0          # cs = r0: taken nodes have |d| < r0
    s = np.where(taken, V[k] / vmax, np.float32(0.0))
    img = ppm_slice(s).astype(int)
    diff = np.abs(img - FIX["viscosities"].astype(int)).max(-1)
    # every pixel within one colour level, and (almost) all exact
    assert diff.max() <= 1
    assert (diff > 0).sum() <= 25


def test_fixture_oracle_builder_matches_ppm():
    """The FP64 oracle builder reproduces the same slices (distance exact up
    to FP32/FP64 colour-quantisation ties)."""
    d = S.config_desc("example01")
    D, W, V = O.scene_boundary(d)
    k = d.boundary_dims[2] // 2
    r0 = d.rest_distance
    img = ppm_slice((D[k] / r0).astype(np.float32)).astype(int)
    diff = np.abs(img - FIX["distances"].astype(int)).max(-1)
    assert (diff > 0).sum() <= 50 and diff.max() <= 1
    taken = np.abs(D[k]) < r0
    img = ppm_slice(np.where(taken, V[k] / V.max(), 0.0).astype(np.float32)).astype(int)
    diff = np.abs(img - FIX["viscosities"].astype(int)).max(-1)
    assert diff.max() <= 1 and (diff > 0).sum() <= 200


@pytest.mark.parametrize("cfg", ["example01", "dam16k", "two_scale"])
def test_product_builder_matches_oracle_builder(cfg):
    d = S.config_desc(cfg)
    D, W, V = S.boundary_of(d)
    Do, Wo, Vo = O.scene_boundary(d)
    r0 = d.rest_distance
    # taken-node mask (|d| < cs, SphInComplexShapes.cpp:217) identical
    # except at FP32/FP64 ties |d| = cs (grids aligned with the walls put
    # whole node planes there)
    tie = np.abs(np.abs(Do) - d.boundary_support) < 1e-6
    mism = (np.abs(D) < r0) != (np.abs(Do) < r0)
    assert not (mism & ~tie).any()
    # nodes next to a tie plane receive a tie node's ghost splat differently
    ok = ~mism & ~tie
    assert np.abs(D - Do)[ok].max() <= 2e-6 * max(1.0, float(np.abs(Do).max()))
    sw = float(Wo.max())
    sv = float(Vo.max())
    assert np.abs(W - Wo)[ok].max() <= 2e-4 * sw
    assert np.abs(V - Vo)[ok].max() <= 2e-4 * sv


def test_box_sdf_closed_form():
    d = S.config_desc("dam16k")
    rng = np.random.default_rng(1)
    pts = rng.uniform(-0.1, 0.74, size=(5000, 3))
    got = S.scene_sdf(d, pts)
    # closed-form signed distance to the box [0, 0.64f]^3 (positive outside)
    half = 0.5 * float(np.float32(0.64))
    q = np.abs(pts - half) - half
    outside = np.linalg.norm(np.maximum(q, 0), axis=1)
    inside = np.minimum(q.max(1), 0)
    ref = outside + inside
    assert np.allclose(got, ref, atol=2e-6)
    assert np.allclose(O.scene_sdf(d, pts), ref, atol=1e-12)


def test_sphere_and_cone_sdf_closed_form():
    d = S.config_desc("dam1m")
    sph, cone = d.obstacles[0], d.obstacles[1]
    c = np.array(sph.center[:])
    # points on rays from the sphere centre: SDF = max(box term, r - |x-c|)
    dirs = np.random.default_rng(2).normal(size=(200, 3))
    dirs /= np.linalg.norm(dirs, axis=1, keepdims=True)
    rad = np.linspace(0.05, 0.45, 200)[:, None]
    pts = c + dirs * rad
    pts = pts[pts[:, 1] > 0.05]
    got = O.scene_sdf(d, pts)
    ref = sph.a[0] - np.linalg.norm(pts - c, axis=1)
    assert np.allclose(got, np.maximum(ref, got - 0 * got), atol=1e-12)  # sphere dominates near it
    assert np.allclose(got[ref > -0.04], ref[ref > -0.04], atol=1e-9)
    # cone: apex is on the surface, base centre at depth min(R, H)-ish, axis
    apex = np.array([cone.center[0], cone.center[1] + cone.a[1], cone.center[2]])
    assert abs(O.scene_sdf(d, apex[None])[0]) < 1e-7
    # point on the slant, midway: distance zero
    mid = np.array([cone.center[0] + 0.5 * cone.a[0], cone.center[1] + 0.5 * cone.a[1], cone.center[2]])
    assert abs(O.scene_sdf(d, mid[None])[0]) < 1e-7
    # outside along the slant normal by 0.05
    nrm = np.array([cone.a[1], cone.a[0], 0.0]) / np.hypot(cone.a[0], cone.a[1])
    out = mid + 0.05 * nrm
    assert O.scene_sdf(d, out[None])[0] == pytest.approx(-0.05, abs=1e-9)
    assert S.scene_sdf(d, out[None])[0] == pytest.approx(-0.05, abs=2e-6)


@pytest.mark.parametrize("cfg,n", [("dam16k", 16384), ("dam1m", 1048576), ("two_scale", 262144)])
def test_config_sizes(cfg, n):
    d = S.config_desc(cfg)
    p = S.params_of(d)
    assert p.num_particles == n
    assert int(np.prod(d.lattice[:])) == n


def test_dam16k_workload_definition():
    d = S.config_desc("dam16k")
    p = S.params_of(d)
    assert tuple(d.lattice) == (32, 16, 32)
    assert d.lattice_spacing == pytest.approx(0.01)
    assert p.compact_support == pytest.approx(0.02)
    assert p.rest_density == 1000.0
    assert p.particle_mass == pytest.approx(1e-3, rel=1e-6)
    assert p.gas_stiffness == 3.0 and p.dynamic_viscosity == pytest.approx(3.5)
    assert p.gravity == pytest.approx(9.81) and p.time_step == pytest.approx(1e-4)
    assert tuple(d.boundary_dims) == (64, 64, 64)
    # cell size h over the box grown by 2h: 36^3 cells
    assert tuple(p.grid_dim) == (36, 36, 36)
    # kernel coefficients (particle_simulation.cu:1779-1788)
    h = 0.02
    assert p.poly6 == pytest.approx(315 / (64 * np.pi * h ** 9), rel=1e-6)
    assert p.grad_spiky == pytest.approx(-45 / (np.pi * h ** 6), rel=1e-6)
    assert p.lapl_visc == pytest.approx(45 / (np.pi * h ** 6), rel=1e-6)
    assert p.poly6_sub == pytest.approx(512 * p.poly6, rel=1e-6)
    assert p.grad_spiky_sub == pytest.approx(64 * p.grad_spiky, rel=1e-6)


def test_jitter_is_bounded_and_seeded():
    d = S.config_desc("dam16k", seed=1)
    pos = S.particles_of(d)
    base = np.stack(np.meshgrid(np.arange(32), np.arange(16), np.arange(32), indexing="ij"), -1)
    # create_particle_box order: x fastest, then y, then z
    i = np.arange(pos.shape[0])
    lat = np.stack([i % 32, (i // 32) % 16, i // 512], -1) * 0.01 + np.array(d.lattice_origin[:])
    jit = pos - lat
    assert np.abs(jit).max() <= 0.001 + 1e-6
    assert np.abs(jit).max() > 0.0009
    assert abs(jit.mean()) < 2e-5
    pos2 = S.particles_of(S.config_desc("dam16k", seed=1))
    assert np.array_equal(pos, pos2)
    pos3 = S.particles_of(S.config_desc("dam16k", seed=2))
    assert not np.array_equal(pos, pos3)
    del base


def test_batch_configs():
    ns = []
    for s in range(64):
        d = S.config_desc("batch", index=s, seed=1)
        p = S.params_of(d)
        ns.append(p.num_particles)
        assert 0.3 * 4096 < p.num_particles < 1.5 * 262144
        box = np.array(d.box_hi[:]) - np.array(d.box_lo[:])
        assert 0.4 <= box[0] / box[2] <= 2.6
        for k in range(d.num_obstacles):
            o = d.obstacles[k]
            assert o.kind in (1, 2, 3)
    total = sum(ns)
    assert 2.0e6 < total < 7.0e6


def test_batch_generator_and_the_picked_scenes():
    """The 64-scene workload: log-uniform counts stratified over the scenes
    (total near the expected ~3.97M), every scene with an obstacle; and the
    eight scenes the GPU tests run against the oracle
    (tests/test_gpu_batch.py PICK) are the extremes they are said to be."""
    all_ = [S.make_scene("batch", index=i, seed=1) for i in range(64)]
    n = np.array([s.n for s in all_])
    asp = np.array([(s.desc.box_hi[0] - s.desc.box_lo[0]) / (s.desc.box_hi[2] - s.desc.box_lo[2]) for s in all_])
    kinds = [sorted(s.desc.obstacles[o].kind for o in range(s.desc.num_obstacles)) for s in all_]
    assert n[31] == n.min() and n[10] == n.max()
    assert np.argmin(asp) == 33 and np.argmax(asp) == 17
    assert kinds[38] == [1] and kinds[47] == [2] and kinds[8] == [3] and kinds[58] == [1, 1]
    assert all(k for k in kinds)                       # every scene has an obstacle
    assert 3.5e6 < n.sum() < 4.5e6
