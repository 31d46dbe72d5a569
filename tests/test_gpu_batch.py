"""Batched scenes (the 64-scene workload: particle counts log-uniform in
[4,096, 262,144], box aspect ratios in [0.5, 2], random sphere / cone / box
obstacles on ~128^3 grids) on the GPU.

Against the FP64 oracle, eight scenes of the batch chosen to cover its
extremes -- the smallest and the largest particle count, each obstacle kind,
the two aspect-ratio extremes, a union of two solids -- run together as one
batch (one launch per stage over all eight):
  * one step: keys, sorted order and cellStart/cellEnd bit-exact per scene,
    density within 1e-5, acceleration within 1e-4 of max(|a|, g);
  * 500 steps (graph replays, the hit-mask kernel pair the autotuner picks
    at this size): per-scene mean density and kinetic energy within 1%;
    with a learned boundary per scene, mean density within 0.1% and kinetic
    energy within 1%.
The full 64-scene batch equals the 64 scenes run as separate handles bit for
bit, and its per-scene sort invariants hold.
"""
import numpy as np
import pytest

import cudasphnew_amd as S
from cudasphnew_amd import Field
from oracle import oracle as O
from tests.helpers import kinetic_energy

pytestmark = pytest.mark.gpu

G = 9.81
THREADS = 16
SEED = 1
# smallest (31: 4,000) and largest (10: 256,000) particle count; sphere (38),
# cone (47) and box (8) alone; aspect 0.523 (33: cone + box) and 1.986 (17:
# sphere + box); two spheres (58)
PICK = (31, 10, 38, 47, 8, 33, 17, 58)
STEPS, CHUNK = 500, 100


def _scenes(idx):
    return [S.make_scene("batch", index=i, seed=SEED) for i in idx]


def _check_accel(acc, orc, tol=1e-4, margin=2e-5):
    acc_o = orc.read(Field.ACCEL)
    scale = np.maximum(np.linalg.norm(acc_o, axis=1), G)[:, None]
    err = (np.abs(acc - acc_o) / scale).max(1)
    on_switch = orc.branch_margin() < margin
    assert on_switch.sum() <= max(2, 1e-3 * acc.shape[0]), on_switch.sum()
    assert err[~on_switch].max() <= tol, err[~on_switch].max()
    return float(err[~on_switch].max())


@pytest.mark.parametrize("tuning", ["default", "hit_masks"])
def test_batch_scenes_one_step_against_oracle(tuning):
    scenes = _scenes(PICK)
    b = S.Batch(scenes)
    if tuning == "hit_masks":
        b.set_tuning(hit_masks=1, hit_list=40)
    b.step(1)
    b.sync()
    for i, sc in zip(range(len(PICK)), scenes):
        keys, flags = O.keys(sc.params, sc.positions, real=1)
        assert flags == 0
        sk, si, cs, ce = O.sort_cells(sc.params, keys)
        assert np.array_equal(b.read(i, Field.SORTED_KEYS), sk), PICK[i]
        assert np.array_equal(b.read(i, Field.SORTED_IDS), si), PICK[i]
        assert np.array_equal(b.read(i, Field.CELL_START), cs), PICK[i]
        assert np.array_equal(b.read(i, Field.CELL_END), ce), PICK[i]
        orc = O.OracleSim(sc, real=0, threads=THREADS)
        orc.step(1)
        rho, rho_o = b.read(i, Field.DENSITY), orc.read(Field.DENSITY)
        er = float(np.max(np.abs(rho - rho_o) / rho_o))
        ea = _check_accel(b.read(i, Field.ACCEL), orc)
        print(f"scene {PICK[i]} ({sc.n}): density {er:.2e}, acceleration {ea:.2e}")
        assert er <= 1e-5


@pytest.fixture(scope="module")
def oracle_trajectories():
    """Mean density and kinetic energy of each picked scene every CHUNK
    steps of the FP64 oracle."""
    out = {}
    for i in PICK:
        sc = S.make_scene("batch", index=i, seed=SEED)
        orc = O.OracleSim(sc, real=0, threads=THREADS)
        m = float(sc.params.particle_mass)
        rows = []
        for _ in range(STEPS // CHUNK):
            orc.step(CHUNK)
            rows.append((orc.read(Field.DENSITY).mean(), kinetic_energy(m, orc.read(Field.VELOCITY))))
        out[i] = rows
        print(f"oracle scene {i} ({sc.n}): {STEPS} steps", flush=True)
    return out


def _run_against(b, scenes, traj, rho_tol, ke_tol):
    worst = [0.0, 0.0]
    for c in range(STEPS // CHUNK):
        b.step(CHUNK)
        b.sync()
        for i, sc in enumerate(scenes):
            m = float(sc.params.particle_mass)
            rho = float(b.read(i, Field.DENSITY).astype(np.float64).mean())
            ke = kinetic_energy(m, b.read(i, Field.VELOCITY))
            rho_o, ke_o = traj[PICK[i]][c]
            dr, dk = abs(rho / rho_o - 1), abs(ke / ke_o - 1)
            worst = [max(worst[0], dr), max(worst[1], dk)]
            assert dr <= rho_tol, (PICK[i], (c + 1) * CHUNK, dr)
            assert dk <= ke_tol, (PICK[i], (c + 1) * CHUNK, dk)
        print(f"step {(c + 1) * CHUNK}: worst mean-density {worst[0]:.2e}, kinetic energy {worst[1]:.2e}",
              flush=True)


def test_batch_scenes_500_steps_against_oracle(oracle_trajectories):
    scenes = _scenes(PICK)
    b = S.Batch(scenes)
    b.set_tuning(hit_masks=1, hit_list=40)
    b.use_graph(True)
    _run_against(b, scenes, oracle_trajectories, 0.01, 0.01)


def test_batch_scenes_with_learned_boundary_500_steps_against_oracle(oracle_trajectories):
    """The batched + learned-boundary form (a network per scene, trained on
    the GPU as the benchmark's batch form trains it): mean density within
    0.1% and kinetic energy within 1% of the FP64 oracle's grid step."""
    from cudasphnew_amd import neural as N
    scenes = _scenes(PICK)
    b = S.Batch(scenes)
    nets = []
    for i, sc in enumerate(scenes):
        net = N.train_for_scene(sc.desc, steps=2000, n_samples=1 << 20)
        rep = b.attach_neural(i, net, val=N.validation_samples(sc.desc, 1 << 18))
        print(f"scene {PICK[i]}: cells network {rep['cells_neural']}, grid {rep['cells_grid']}; bricks "
              f"{rep['bricks_checked']} checked, {rep['bricks_fallback']} on the grid "
              f"({rep['bricks_fallback_normal']} for the normal); mean error {rep['mean_error'] / sc.desc.h:.3f} h, "
              f"normal {rep['mean_normal_error']:.3f}", flush=True)
        assert rep["cells_neural"] > 0
        nets.append(net)
    b.set_tuning(hit_masks=1, hit_list=40)
    b.use_graph(True)
    _run_against(b, scenes, oracle_trajectories, 1e-3, 0.01)
    del nets


def test_full_batch_matches_handles_and_sorts():
    """All 64 scenes (about 4M particles) as one batch, 3 steps: every scene
    equals its own handle bit for bit (positions, velocities, densities,
    sorted ids and keys), and each scene's sorted keys are non-decreasing
    with ties in original-id order and cell ranges consistent."""
    scenes = _scenes(range(64))
    b = S.Batch(scenes)
    b.use_graph(True)
    b.step(3)
    b.sync()
    for i, sc in enumerate(scenes):
        sim = S.ParticleSimulation(sc).Init()
        sim.Advance(3)
        sim.sync()
        for f in (Field.POSITION, Field.VELOCITY, Field.DENSITY, Field.SORTED_IDS, Field.SORTED_KEYS):
            assert np.array_equal(b.read(i, f), sim.read(f)), (i, f)
        sim.close()
        sk, si = b.read(i, Field.SORTED_KEYS), b.read(i, Field.SORTED_IDS)
        assert np.all(np.diff(sk) >= 0)
        assert np.all(np.diff(si)[np.diff(sk) == 0] > 0)
        assert np.array_equal(np.sort(si), np.arange(sc.n))
        cs, ce = b.read(i, Field.CELL_START), b.read(i, Field.CELL_END)
        occ = cs >= 0
        assert np.array_equal(np.repeat(np.nonzero(occ)[0], (ce - cs)[occ]), sk)


@pytest.mark.parametrize("batch_tuning", [dict(hit_list=40, block=128), dict(hit_list=48, block=256)],
                         ids=lambda t: f"{t['hit_list']}-{t['block']}")
def test_batch_hit_mask_work_list_matches_handles(batch_tuning):
    """The batch's hit-mask pair runs one wavefront per work-list entry, none
    across a scene boundary: the 64-slot-aligned wavefronts, then the lanes
    of each scene that starts inside one.  Each particle's arithmetic is the
    handle's, so every scene equals its own handle bit for bit over 3 steps
    -- here the eight picked scenes plus two of fewer than 64 particles (a
    scene inside one wavefront) placed between them."""
    picked = _scenes(PICK)
    tiny = [picked[0].with_positions(picked[0].positions[:30]), picked[3].with_positions(picked[3].positions[:7])]
    scenes = picked[:2] + [tiny[0]] + picked[2:5] + [tiny[1]] + picked[5:]
    b = S.Batch(scenes)
    b.set_tuning(hit_masks=1, **batch_tuning)
    b.step(3)
    b.sync()
    for i, sc in enumerate(scenes):
        sim = S.ParticleSimulation(sc).Init()
        sim.set_tuning(hit_masks=1, hit_list=40)
        sim.Advance(3)
        sim.sync()
        for f in (Field.POSITION, Field.VELOCITY, Field.DENSITY, Field.ACCEL, Field.SORTED_IDS):
            assert np.array_equal(b.read(i, f), sim.read(f)), (i, sc.n, f)
        sim.close()
