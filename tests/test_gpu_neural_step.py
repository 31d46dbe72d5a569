"""The SPH step with the learned boundary attached, against the grid-based
FP64 oracle and against the grid form of the same step.

The network (trained on the distance texture) supplies the wall normals of
the penalty and the collision response in its cells; cells whose texels all
hold the far-field values take no wall term.  Bars: far and grid cells keep
the grid form's arithmetic exactly; over 1000 steps the mean density stays
within 0.1% of the FP64 oracle and the kinetic energy within 1%; the
network's RMS error over the validation band is below 0.05 h."""
import numpy as np
import pytest

import cudasphnew_amd as S
from cudasphnew_amd import Field
from cudasphnew_amd import neural as N
from oracle import oracle as O
from tests.helpers import kinetic_energy

pytestmark = pytest.mark.gpu
THREADS = 16


@pytest.fixture(scope="module")
def dam16k_net():
    sc = S.make_scene("dam16k")
    net = N.train_for_scene(sc.desc)
    return sc, net


def test_attach_report(dam16k_net):
    sc, net = dam16k_net
    sim = S.ParticleSimulation(sc).Init()
    rep = sim.attach_neural(net)
    print(rep)
    ncell = int(np.prod([sc.params.grid_dim[a] for a in range(3)]))
    assert rep["cells_none"] + rep["cells_neural"] + rep["cells_grid"] == ncell
    assert rep["cells_neural"] > 0 and rep["cells_none"] > 0
    assert rep["samples_in_band"] > 10000
    assert rep["mean_error"] < 0.05 * sc.desc.h
    # bricks fall back where the distance (0.05 h) or the penetration-weighted
    # normal (0.05) misses its bar; the network serves most wall cells
    assert rep["bricks_fallback"] <= 0.15 * rep["bricks_checked"]
    assert rep["cells_neural"] > 2 * rep["cells_grid"]
    assert rep["mean_normal_error"] < 0.05


def test_attach_releases_the_distance_copy(dam16k_net):
    """With a network on every scene the float copy of the distance grid is
    freed (4 bytes per node; the cells that fall back to the grid read the
    float4 grid's distance channel); detaching allocates it again from that
    channel, and the plain step is unchanged."""
    sc, net = dam16k_net
    nodes = int(np.prod(sc.boundary_dims))
    n = sc.params.num_particles
    a = S.ParticleSimulation(sc).Init()
    b = S.ParticleSimulation(sc).Init()
    b0 = a.GetSizeMemoryGPU()
    a.attach_neural(net)
    b1 = a.GetSizeMemoryGPU()
    print(f"device bytes: grid form {b0}, learned boundary {b1}; {nodes} nodes, {n} particles")
    assert b0 - b1 == 4 * nodes - 21 * n   # distance copy out, per-particle records (16 + 1 + 4 B) in
    a.Advance(5)
    a.detach_neural()
    assert a.GetSizeMemoryGPU() == b0
    c = S.ParticleSimulation(sc).Init()
    c.attach_neural(net)
    c.detach_neural()
    c.Advance(20)
    b.Advance(20)
    c.sync()
    b.sync()
    for f in (Field.POSITION, Field.VELOCITY, Field.DENSITY):
        assert np.array_equal(c.read(f), b.read(f))


def test_far_and_grid_cells_follow_the_grid_form(dam16k_net):
    """Lattice moved onto the floor so that the wall terms act at once."""
    sc, net = dam16k_net
    pos = sc.positions.copy()
    pos[:, 1] -= 0.025
    sc2 = sc.with_positions(pos)
    a = S.ParticleSimulation(sc2).Init()
    a.attach_neural(net)
    b = S.ParticleSimulation(sc2).Init()
    a.Advance(1)
    b.Advance(1)
    a.sync()
    b.sync()
    src = a.read(Field.WALL_SOURCE)
    print("sources (none, network, grid):", np.bincount(src, minlength=3))
    assert (src == 1).sum() > 1000 and (src == 0).sum() > 1000
    # densities: wall density comes from the grid in every non-far cell
    assert np.array_equal(a.read(Field.DENSITY), b.read(Field.DENSITY))
    acc_n, acc_g = a.read(Field.ACCEL), b.read(Field.ACCEL)
    far = src == 0
    assert np.array_equal(acc_n[far], acc_g[far])
    # network normals: the grid's central differences span two spacings
    # across the clamp of the distance texture, the network's gradient is
    # local, so magnitudes differ near the clamp; both push off the floor
    pen = np.linalg.norm(acc_g, axis=1) > 1e3
    rel = np.linalg.norm(acc_n[pen] - acc_g[pen], axis=1) / np.linalg.norm(acc_g[pen], axis=1)
    cos = np.sum(acc_n[pen] * acc_g[pen], 1) / (np.linalg.norm(acc_n[pen], axis=1) * np.linalg.norm(acc_g[pen], axis=1))
    print(f"penalty particles {pen.sum()}: relative acceleration difference median {np.median(rel):.3g}, "
          f"max {rel.max():.3g}; cosine min {cos.min():.4f}")
    assert np.all(acc_n[pen, 1] > 0) and np.all(acc_g[pen, 1] > 0)
    assert cos.min() > 0.99


def test_neural_run_mean_density_and_energy(dam16k_net):
    sc, net = dam16k_net
    sim = S.ParticleSimulation(sc).Init()
    sim.attach_neural(net)
    orc = O.OracleSim(sc, real=0, threads=THREADS)
    m = sc.params.particle_mass
    for chunk in range(4):
        sim.Advance(250)
        orc.step(250)
        sim.sync()
        rho, rho_o = sim.read(Field.DENSITY), orc.read(Field.DENSITY)
        ke = kinetic_energy(m, sim.read(Field.VELOCITY))
        ke_o = kinetic_energy(m, orc.read(Field.VELOCITY))
        src = np.bincount(sim.read(Field.WALL_SOURCE), minlength=3)
        print(f"step {250 * (chunk + 1)}: mean rho {rho.mean():.5f} vs {rho_o.mean():.5f} "
              f"({abs(rho.mean() / rho_o.mean() - 1):.2e}), KE {ke:.5g} vs {ke_o:.5g}, sources {src}")
        assert abs(rho.mean() - rho_o.mean()) <= 1e-3 * rho_o.mean()
        assert abs(ke - ke_o) <= 0.01 * ke_o
    assert src[1] > 0.2 * sc.n


def test_detach_restores_grid_form(dam16k_net):
    sc, net = dam16k_net
    a = S.ParticleSimulation(sc).Init()
    a.attach_neural(net)
    a.detach_neural()
    a.Advance(3)
    b = S.ParticleSimulation(sc).Init()
    b.Advance(3)
    a.sync()
    b.sync()
    assert np.array_equal(a.read(Field.POSITION), b.read(Field.POSITION))


def test_network_is_a_compact_stand_in_for_the_1m_distance_grid():
    """What the network gains over the grid: at the 1M-particle scene it
    holds the distance texture in 16,641 parameters (66.6 KB in FP32) in
    place of 256x128x256 nodes (33.5 MB), with its validation RMS distance
    error within 0.05 h and its normal within the penetration-weighted bar
    in most bricks, and the step that uses its normals keeps mean density and
    kinetic energy with the grid form."""
    sc = S.make_scene("dam1m")
    h = float(sc.desc.h)
    net = N.train_for_scene(sc.desc)
    grid_bytes = 4 * int(np.prod(sc.boundary_dims))
    net_bytes = 4 * N.NUM_PARAMS
    sim = S.ParticleSimulation(sc).Init()
    rep = sim.attach_neural(net)
    print(f"grid {grid_bytes / 1e6:.1f} MB vs network {net_bytes / 1e3:.1f} KB ({grid_bytes / net_bytes:.0f}x); {rep}")
    assert grid_bytes > 400 * net_bytes
    assert rep["mean_error"] < 0.05 * h
    assert rep["bricks_fallback"] <= 0.15 * rep["bricks_checked"]
    assert rep["cells_neural"] > 2 * rep["cells_grid"]
    assert rep["mean_normal_error"] < 0.05
    ref = S.ParticleSimulation(sc).Init()
    sim.Advance(200)
    ref.Advance(200)
    sim.sync()
    ref.sync()
    m = sc.params.particle_mass
    rho, rho_g = sim.read(Field.DENSITY), ref.read(Field.DENSITY)
    ke, ke_g = kinetic_energy(m, sim.read(Field.VELOCITY)), kinetic_energy(m, ref.read(Field.VELOCITY))
    print(f"200 steps: mean rho {rho.mean():.5f} vs grid form {rho_g.mean():.5f}, KE {ke:.5g} vs {ke_g:.5g}")
    assert abs(rho.mean() / rho_g.mean() - 1) <= 1e-3
    assert abs(ke / ke_g - 1) <= 0.01
