"""Render interop view (ParticleSimulation::map, particle_simulation.cu:2685-2700):
the mapped device buffers hold exactly what the read-back API returns.

Positions by original id carry (x, y, z) of the last step and its density;
sub-particle positions are indexed by sub id; the default-state id list and
the active sub-particle id list (regular first, then boundary) are the ones a
renderer draws from (twoscalestatesub.cpp:96-113).  All integer or copied
data: bit-exact."""
import ctypes as C

import numpy as np
import pytest

import cudasphnew_amd as S
from cudasphnew_amd import Field

pytestmark = pytest.mark.gpu


def _hip():
    return C.CDLL("libamdhip64.so")


def _fetch(ptr, count, dtype):
    out = np.empty(count, dtype=dtype)
    if count:
        hip = _hip()
        hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        assert hip.hipMemcpy(out.ctypes.data, ptr, out.nbytes, 2) == 0   # device to host
    return out


@pytest.fixture(scope="module")
def dam16k():
    return S.make_scene("dam16k")


def test_plain_step_view(dam16k):
    sim = S.ParticleSimulation(dam16k).Init()
    sim.Advance(3)
    v = sim.map()
    n = dam16k.n
    assert v["num_particles"] == n and v["positions"]
    assert not v["sub_positions"] and not v["default_ids"] and not v["sub_ids"]
    assert v["num_default"] == v["num_sub_regular"] == v["num_sub_boundary"] == 0
    p = _fetch(v["positions"], 4 * n, np.float32).reshape(n, 4)
    assert np.array_equal(p[:, :3], sim.read(Field.POSITION))
    assert np.array_equal(p[:, 3], sim.read(Field.DENSITY))
    sim.unmap()
    # the view follows the state: one more step, map again
    sim.Advance(1)
    p2 = _fetch(sim.map()["positions"], 4 * n, np.float32).reshape(n, 4)
    assert np.array_equal(p2[:, :3], sim.read(Field.POSITION))
    assert not np.array_equal(p2[:, :3], p[:, :3])


def test_two_scale_view(dam16k):
    sc = dam16k.with_positions(dam16k.positions)
    sim = S.ParticleSimulation(sc).Init()
    sim.AdvanceTwoScale(2)
    v = sim.map()
    c = sim.counts()
    n = sc.n
    assert v["num_default"] == c.num_default
    assert v["num_sub_regular"] == 8 * c.num_split and v["num_sub_boundary"] == 8 * c.num_boundary
    assert v["num_sub_regular"] + v["num_sub_boundary"] == c.num_sub > 0
    p = _fetch(v["positions"], 4 * n, np.float32).reshape(n, 4)
    assert np.array_equal(p[:, :3], sim.read(Field.POSITION))
    ids = _fetch(v["default_ids"], v["num_default"], np.uint32)
    assert np.array_equal(ids, sim.read(Field.DEFAULT_IDS))
    sub_ids = _fetch(v["sub_ids"], c.num_sub, np.uint32)
    assert np.array_equal(sub_ids, sim.read(Field.SUB_IDS))
    sp = _fetch(v["sub_positions"], 4 * 8 * n, np.float32).reshape(8 * n, 4)
    assert np.array_equal(sp[sub_ids, :3], sim.read(Field.SUB_POSITION))
    assert np.array_equal(sp[sub_ids, 3], sim.read(Field.SUB_DENSITY))
    # a default particle is drawn from the coarse buffer, a split one is not
    split = sim.read(Field.SPLIT_IDS)
    assert not np.isin(split, ids).any()
    # children of a split parent: sub ids 8 * parent + child, regular first
    reg = sub_ids[:v["num_sub_regular"]]
    assert np.array_equal(np.unique(reg // 8), np.sort(split))
