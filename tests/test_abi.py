"""The C ABI library loads, exports every declared symbol, and its structs
have the layout of include/sph.h.  No GPU needed."""
import ctypes as C
import os
import re
import subprocess

import numpy as np
import pytest

import cudasphnew_amd as S
from cudasphnew_amd import _native as N
from oracle import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "sph.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(sph_[a-z0-9_]+)\s*\(", text)))


def test_every_declared_symbol_is_exported():
    names = declared_functions()
    assert len(names) >= 30
    nm = subprocess.run(["nm", "-D", "--defined-only", N.LIB_PATH], capture_output=True, text=True).stdout
    exported = set(re.findall(r"\bT (sph_\w+)", nm))
    missing = [n for n in names if n not in exported]
    assert not missing, missing
    for n in names:
        assert hasattr(N.lib, n)


def test_bindings_cover_header():
    bound = {s[0] for s in N.SIGNATURES}
    assert set(declared_functions()) == bound


def test_struct_layout_matches_header():
    lay = O.struct_layout()
    mine = [C.sizeof(N.SphParams), C.sizeof(N.SphBoundary), C.sizeof(N.SphScene), C.sizeof(N.SceneDesc),
            C.sizeof(N.SphCounts), C.sizeof(N.SphSolid), N.SphParams.max_sub_particles.offset,
            N.SphScene.boundary.offset, N.SceneDesc.sub_placement.offset, N.SphBoundary.viscosity.offset,
            C.sizeof(N.SphMapSolid), N.SphMapSolid.distance.offset, N.SphMapSolid.solid_hi.offset]
    assert lay == mine


def test_abi_version_and_names():
    assert N.lib.sph_abi_version() == 1
    assert N.lib.sph_stage_name(6) == b"force"
    assert N.lib.sph_stage_name(99) == b""


def test_create_rejects_bad_arguments_without_touching_device():
    sc = S.make_scene("dam16k")
    h = C.c_void_p()
    # NULL params
    assert N.lib.sph_create(None, None, C.byref(h)) == N.Status.INVALID_ARG
    # particle count mismatch
    p = N.SphParams.from_buffer_copy(sc.params)
    p.num_particles = 7
    s = sc.sph_scene()
    assert N.lib.sph_create(C.byref(p), C.byref(s), C.byref(h)) == N.Status.INVALID_ARG
    assert b"num_particles" in N.lib.sph_last_error_message(None)


@pytest.mark.skipif(S.device_count() > 0, reason="checks the no-device path")
def test_create_without_device_reports_no_device():
    sc = S.make_scene("dam16k")
    h = C.c_void_p()
    s = sc.sph_scene()
    st = N.lib.sph_create(C.byref(sc.params), C.byref(s), C.byref(h))
    assert st == N.Status.NO_DEVICE
    with pytest.raises(S.SphError):
        S.ParticleSimulation(sc).Init()


def test_null_handle_calls_are_rejected():
    assert N.lib.sph_step(None, 1) == N.Status.INVALID_ARG
    assert N.lib.sph_sync(None) == N.Status.INVALID_ARG
    assert N.lib.sph_destroy(None) == N.Status.INVALID_ARG
    assert N.lib.sph_batch_step(None, 1) == N.Status.INVALID_ARG
    out = np.zeros(3, np.float32)
    assert N.lib.sph_read_positions(None, N.as_fptr(out), 1) == N.Status.INVALID_ARG
