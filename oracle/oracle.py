"""ctypes wrapper of the CPU oracle liboracle_sph.so -- TEST INFRASTRUCTURE.

Only tests/, __graft_entry__.smoke() and the cpu_baseline leg of bench.py
use this module, as the checker; the product never imports it.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from cudasphnew_amd._native import Field, SceneDesc, SphBoundary, SphCounts, SphParams

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle_sph.so")

FLAG_OUT_OF_DOMAIN = 1
FLAG_NONFINITE = 2
FIELD_COLOR_GRAD_NORM = 100
FIELD_WALL_DISTANCE = 101
FIELD_SUB_COLOR_GRAD_NORM = 102


def _load():
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} missing; run `make -C oracle`")
    L = C.CDLL(LIB_PATH)
    vp, dp, fp, ip = C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_float), C.POINTER(C.c_int32)
    sig = {
        "orc_create": (vp, [C.POINTER(SphParams), C.POINTER(SphBoundary), C.c_uint32, fp, fp, C.c_int, C.c_int]),
        "orc_create_f64": (vp, [C.POINTER(SphParams), C.POINTER(SphBoundary), C.c_uint32, dp, dp, C.c_int,
                                C.c_int]),
        "orc_destroy": (None, [vp]),
        "orc_step": (C.c_int, [vp, C.c_int, C.c_int]),
        "orc_step_sub": (C.c_int, [vp, C.c_int]),
        "orc_read": (C.c_long, [vp, C.c_int, vp, C.c_size_t]),
        "orc_counts": (None, [vp, C.POINTER(SphCounts)]),
        "orc_flags": (C.c_uint32, [vp]),
        "orc_last_step_seconds": (C.c_double, [vp]),
        "orc_keys": (C.c_uint32, [C.POINTER(SphParams), C.c_uint32, fp, ip, C.c_int]),
        "orc_sort_cells": (None, [C.POINTER(SphParams), C.c_uint32, ip, ip, ip, ip, ip]),
        "orc_tex": (None, [C.POINTER(SphBoundary), C.c_int, C.c_uint32, dp, dp]),
        "orc_density_brute": (None, [C.POINTER(SphParams), C.POINTER(SphBoundary), C.c_uint32, dp, dp]),
        "orc_accel_brute": (None, [C.POINTER(SphParams), C.POINTER(SphBoundary), C.c_uint32, dp, dp, dp, dp]),
        "orc_sub_forces": (C.c_int, [C.POINTER(SphParams), C.POINTER(SphBoundary), C.c_uint32, C.c_uint32, dp, dp,
                                     dp, C.c_int, dp, dp, dp]),
        "orc_scene_boundary": (C.c_int, [C.POINTER(SceneDesc), dp, dp, dp]),
        "orc_scene_sdf": (None, [C.POINTER(SceneDesc), C.c_uint32, dp, dp]),
        "orc_struct_layout": (None, [C.POINTER(C.c_size_t)]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    return L


lib = _load()


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _fp(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def _ip(a):
    return a.ctypes.data_as(C.POINTER(C.c_int32))


class OracleSim:
    """Stateful oracle simulation in original-id order.  real=0: FP64
    oracle, real=1: FP32 transcription (CPU baseline)."""

    def __init__(self, scene, real: int = 0, threads: int = 1, positions=None, velocities=None):
        self.scene = scene
        pos = scene.positions if positions is None else positions
        vel = scene.velocities if velocities is None else velocities
        self._b = scene.sph_boundary()
        self._p = SphParams.from_buffer_copy(scene.params)
        self._p.num_particles = pos.shape[0]
        self.n = pos.shape[0]
        if pos.dtype == np.float64:
            pd = np.ascontiguousarray(pos, np.float64)
            vd = None if vel is None else np.ascontiguousarray(vel, np.float64)
            self._s = lib.orc_create_f64(C.byref(self._p), C.byref(self._b), self.n, _dp(pd),
                                         None if vd is None else _dp(vd), real, threads)
        else:
            pf = np.ascontiguousarray(pos, np.float32)
            vf = None if vel is None else np.ascontiguousarray(vel, np.float32)
            self._s = lib.orc_create(C.byref(self._p), C.byref(self._b), self.n, _fp(pf),
                                     None if vf is None else _fp(vf), real, threads)
        if not self._s:
            raise RuntimeError("orc_create failed")

    def __del__(self):
        if getattr(self, "_s", None):
            lib.orc_destroy(self._s)
            self._s = None

    def step(self, n: int = 1, two_scale: bool = False) -> float:
        lib.orc_step(self._s, n, 1 if two_scale else 0)
        return float(lib.orc_last_step_seconds(self._s))

    def step_sub(self, n: int = 1) -> None:
        """AdvanceSubParticles: the current sub-particles alone."""
        lib.orc_step_sub(self._s, n)

    def counts(self) -> SphCounts:
        c = SphCounts()
        lib.orc_counts(self._s, C.byref(c))
        return c

    def flags(self) -> int:
        return int(lib.orc_flags(self._s))

    def branch_margin(self) -> np.ndarray:
        """Relative distance of each particle from the two switches of the
        acceleration (|grad c| vs normThresh, wall distance vs r0)."""
        out = []
        for f in (FIELD_COLOR_GRAD_NORM, FIELD_WALL_DISTANCE):
            a = np.empty(self.n, np.float64)
            if lib.orc_read(self._s, f, a.ctypes.data_as(C.c_void_p), a.nbytes) < 0:
                raise RuntimeError("orc_read failed")
            out.append(a)
        cg, dw = out
        thr = float(self._p.norm_thresh)
        r0 = float(self._b.rest_distance)
        # distWall == r0 exactly (constant clamped field, interpolated
        # exactly) is not ambiguous: both precisions skip the wall term
        wall = np.where(dw == r0, np.inf, np.abs(dw / r0 - 1.0))
        return np.minimum(np.abs(cg / thr - 1.0), wall)

    def sub_branch_margin(self) -> np.ndarray:
        """Relative distance of each sub-particle's colour-gradient norm from
        normThresh (inf for sub-particles without an acceleration pass)."""
        n = self.counts().num_sub
        a = np.empty(n, np.float64)
        if lib.orc_read(self._s, FIELD_SUB_COLOR_GRAD_NORM, a.ctypes.data_as(C.c_void_p), a.nbytes) < 0:
            raise RuntimeError("orc_read failed")
        return np.abs(a / float(self._p.norm_thresh) - 1.0)

    def read(self, field) -> np.ndarray:
        field = Field(field)
        c = self.counts()
        n, nsub = self.n, c.num_sub
        ncells = int(np.prod([self._p.grid_dim[a] for a in range(3)]))
        nsubcells = int(np.prod([self._p.grid_dim_sub[a] for a in range(3)]))
        shapes = {
            Field.POSITION: ((n, 3), np.float64), Field.VELOCITY: ((n, 3), np.float64),
            Field.ACCEL: ((n, 3), np.float64), Field.DENSITY: ((n,), np.float64),
            Field.PRESSURE: ((n,), np.float64), Field.STATE: ((n,), np.uint8),
            Field.SORTED_KEYS: ((n,), np.int32), Field.SORTED_IDS: ((n,), np.int32),
            Field.CELL_START: ((ncells,), np.int32), Field.CELL_END: ((ncells,), np.int32),
            Field.SUB_POSITION: ((nsub, 3), np.float64), Field.SUB_VELOCITY: ((nsub, 3), np.float64),
            Field.SUB_ACCEL: ((nsub, 3), np.float64), Field.SUB_DENSITY: ((nsub,), np.float64),
            Field.SUB_PRESSURE: ((nsub,), np.float64), Field.SUB_IDS: ((nsub,), np.int32),
            Field.SPLIT_IDS: ((c.num_split,), np.int32), Field.BOUNDARY_IDS: ((c.num_boundary,), np.int32),
            Field.DEFAULT_IDS: ((c.num_default,), np.int32), Field.SUB_SORTED_KEYS: ((nsub,), np.int32),
            Field.SUB_SORTED_IDS: ((nsub,), np.int32),
            Field.SUB_CELL_START: ((nsubcells,), np.int32), Field.SUB_CELL_END: ((nsubcells,), np.int32),
        }
        shape, dt = shapes[field]
        out = np.empty(shape, dt)
        got = lib.orc_read(self._s, int(field), out.ctypes.data_as(C.c_void_p), out.nbytes)
        if got < 0:
            raise RuntimeError(f"orc_read({field.name}) failed")
        return out


def keys(params: SphParams, pos: np.ndarray, real: int = 1):
    pf = np.ascontiguousarray(pos, np.float32)
    out = np.empty(pf.shape[0], np.int32)
    flags = lib.orc_keys(C.byref(params), pf.shape[0], _fp(pf), _ip(out), real)
    return out, int(flags)


def sort_cells(params: SphParams, keys_: np.ndarray):
    k = np.ascontiguousarray(keys_, np.int32)
    n = k.shape[0]
    nc = int(np.prod([params.grid_dim[a] for a in range(3)]))
    sk, si = np.empty(n, np.int32), np.empty(n, np.int32)
    cs, ce = np.empty(nc, np.int32), np.empty(nc, np.int32)
    lib.orc_sort_cells(C.byref(params), n, _ip(k), _ip(sk), _ip(si), _ip(cs), _ip(ce))
    return sk, si, cs, ce


def tex(scene, which: int, xyz: np.ndarray) -> np.ndarray:
    b = scene.sph_boundary()
    p = np.ascontiguousarray(xyz, np.float64).reshape(-1, 3)
    out = np.empty(p.shape[0], np.float64)
    lib.orc_tex(C.byref(b), which, p.shape[0], _dp(p), _dp(out))
    return out


def density_brute(scene, pos: np.ndarray) -> np.ndarray:
    b = scene.sph_boundary()
    p = np.ascontiguousarray(pos, np.float64)
    out = np.empty(p.shape[0], np.float64)
    lib.orc_density_brute(C.byref(scene.params), C.byref(b), p.shape[0], _dp(p), _dp(out))
    return out


def accel_brute(scene, pos, vel, rho) -> np.ndarray:
    b = scene.sph_boundary()
    p = np.ascontiguousarray(pos, np.float64)
    v = np.ascontiguousarray(vel, np.float64)
    r = np.ascontiguousarray(rho, np.float64)
    out = np.empty_like(p)
    lib.orc_accel_brute(C.byref(scene.params), C.byref(b), p.shape[0], _dp(p), _dp(v), _dp(r), _dp(out))
    return out


def sub_forces(scene, pos, vel, nsplit_subs: int, rho_boundary, threads: int = 8):
    """FP64 sub-particle density and acceleration on a given state (sub-list
    order: split children first, then boundary children with the given
    projected densities).  Returns (rho, acc, colour-gradient norm)."""
    pos = np.ascontiguousarray(pos, np.float64)
    vel = np.ascontiguousarray(vel, np.float64)
    n = pos.shape[0]
    rin = np.zeros(n, np.float64)
    rin[nsplit_subs:] = rho_boundary
    rho = np.empty(n, np.float64)
    acc = np.empty((n, 3), np.float64)
    cgn = np.empty(n, np.float64)
    b = scene.sph_boundary()
    if lib.orc_sub_forces(C.byref(scene.params), C.byref(b), n, nsplit_subs, _dp(pos), _dp(vel), _dp(rin), threads,
                          _dp(rho), _dp(acc), _dp(cgn)) < 0:
        raise ValueError("orc_sub_forces: bad sizes")
    return rho, acc, cgn


def scene_boundary(desc: SceneDesc):
    nx, ny, nz = (int(v) for v in desc.boundary_dims)
    out = [np.empty((nz, ny, nx), np.float64) for _ in range(3)]
    lib.orc_scene_boundary(C.byref(desc), *(_dp(a) for a in out))
    return out


def scene_sdf(desc: SceneDesc, xyz: np.ndarray) -> np.ndarray:
    p = np.ascontiguousarray(xyz, np.float64).reshape(-1, 3)
    out = np.empty(p.shape[0], np.float64)
    lib.orc_scene_sdf(C.byref(desc), p.shape[0], _dp(p), _dp(out))
    return out


def struct_layout():
    out = (C.c_size_t * 13)()
    lib.orc_struct_layout(out)
    return list(out)


def _mesh_sig():
    u8p = C.POINTER(C.c_uint8)
    u32p = C.POINTER(C.c_uint32)
    lib.orc_mesh_sdf.restype = None
    lib.orc_mesh_sdf.argtypes = [C.POINTER(C.c_float), C.c_uint32, u32p, C.c_uint32, C.c_float, C.c_float,
                                 C.c_float, C.POINTER(C.c_int32), C.POINTER(C.c_float), C.POINTER(C.c_float),
                                 C.POINTER(C.c_double), C.POINTER(C.c_double), u8p]
    lib.orc_point_triangle.restype = None
    lib.orc_point_triangle.argtypes = [C.POINTER(C.c_double), C.c_size_t, C.POINTER(C.c_double),
                                       C.POINTER(C.c_double)]
    lib.orc_inside.restype = None
    lib.orc_inside.argtypes = [C.POINTER(C.c_float), u32p, C.c_uint32, C.POINTER(C.c_float), C.c_size_t, u8p]


_mesh_sig()


def mesh_sdf(vertices, faces, rest: float, support: float, dx: float, dims, lo, hi):
    """FP64 signed distance field of a mesh on the given domain: distance,
    normal, present arrays shaped (nz, ny, nx)."""
    v = np.ascontiguousarray(vertices, np.float32)
    f = np.ascontiguousarray(faces, np.uint32)
    shape = (dims[2], dims[1], dims[0])
    d = np.empty(shape, np.float64)
    n = np.empty(shape + (3,), np.float64)
    p = np.empty(shape, np.uint8)
    lib.orc_mesh_sdf(_fp(v), v.shape[0], f.ctypes.data_as(C.POINTER(C.c_uint32)), f.shape[0], float(dx),
                     float(rest), float(support), (C.c_int32 * 3)(*dims), (C.c_float * 3)(*lo),
                     (C.c_float * 3)(*hi), _dp(d), _dp(n), p.ctypes.data_as(C.POINTER(C.c_uint8)))
    return d, n, p


def point_triangle(points, tri):
    p = np.ascontiguousarray(points, np.float64).reshape(-1, 3)
    t = np.ascontiguousarray(tri, np.float64).reshape(9)
    out = np.empty(p.shape[0])
    lib.orc_point_triangle(_dp(p), p.shape[0], _dp(t), _dp(out))
    return out


def inside(vertices, faces, points):
    v = np.ascontiguousarray(vertices, np.float32)
    f = np.ascontiguousarray(faces, np.uint32)
    p = np.ascontiguousarray(points, np.float32).reshape(-1, 3)
    out = np.empty(p.shape[0], np.uint8)
    lib.orc_inside(_fp(v), f.ctypes.data_as(C.POINTER(C.c_uint32)), f.shape[0], _fp(p), p.shape[0],
                   out.ctypes.data_as(C.POINTER(C.c_uint8)))
    return out.astype(bool)
