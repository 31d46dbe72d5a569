"""Mesh -> signed distance field (BoundaryMap of the reference), generated on
the GPU by csrc/mesh.hip; wrappers of the sph_obj_* / sph_mesh_* /
sph_boundary_map_* calls of include/sph.h."""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._native import SphBoundaryMapConfig, SphError, SphMesh, Status, as_fptr, lib


def _check(st, what):
    if st != Status.OK:
        raise SphError(st, what)


def read_obj(path: str):
    """(vertices (n, 3) float32, faces (m, 3) uint32, 0-based)."""
    nv, nf = C.c_uint32(0), C.c_uint32(0)
    _check(lib.sph_obj_read(path.encode(), None, C.byref(nv), None, C.byref(nf)), f"reading {path}")
    v = np.empty((nv.value, 3), np.float32)
    f = np.empty((nf.value, 3), np.uint32)
    _check(lib.sph_obj_read(path.encode(), as_fptr(v), C.byref(nv), f.ctypes.data_as(C.POINTER(C.c_uint32)),
                            C.byref(nf)), f"reading {path}")
    return v, f


def fit(vertices: np.ndarray, lo, hi) -> np.ndarray:
    v = np.ascontiguousarray(vertices, np.float32).copy()
    _check(lib.sph_mesh_fit(as_fptr(v), v.shape[0], (C.c_float * 3)(*lo), (C.c_float * 3)(*hi)), "fit")
    return v


def _mesh(v, f):
    v = np.ascontiguousarray(v, np.float32)
    f = np.ascontiguousarray(f, np.uint32)
    m = SphMesh(as_fptr(v), v.shape[0], f.ctypes.data_as(C.POINTER(C.c_uint32)), f.shape[0])
    return m, (v, f)


def config(rest_distance: float, compact_support: float, dx: float) -> SphBoundaryMapConfig:
    return SphBoundaryMapConfig(float(rest_distance), float(compact_support), float(dx))


def domain(vertices, faces, dx: float):
    m, keep = _mesh(vertices, faces)
    dims, lo, hi = (C.c_int32 * 3)(), (C.c_float * 3)(), (C.c_float * 3)()
    _check(lib.sph_boundary_map_domain(C.byref(m), float(dx), dims, lo, hi), "domain")
    return tuple(dims), np.array(lo[:], np.float32), np.array(hi[:], np.float32)


def generate(vertices, faces, cfg: SphBoundaryMapConfig):
    """dict: dims, lo, hi, distance (nz, ny, nx), normal (nz, ny, nx, 3),
    present (nz, ny, nx) uint8, ms (device time)."""
    m, keep = _mesh(vertices, faces)
    dims, lo, hi = domain(vertices, faces, cfg.dx)
    shape = (dims[2], dims[1], dims[0])
    d = np.empty(shape, np.float32)
    n = np.empty(shape + (3,), np.float32)
    p = np.empty(shape, np.uint8)
    ms = C.c_float()
    _check(lib.sph_boundary_map_generate(C.byref(m), C.byref(cfg), as_fptr(d), as_fptr(n),
                                         p.ctypes.data_as(C.POINTER(C.c_uint8)), C.byref(ms)), "generate")
    return {"dims": dims, "lo": lo, "hi": hi, "distance": d, "normal": n, "present": p, "ms": float(ms.value)}


def grown_domain(vertices, dx: float, margin: float):
    """(dims, lo) of a grid of spacing dx over the mesh's bounding box grown
    by margin on every side."""
    v = np.ascontiguousarray(vertices, np.float32)
    vlo, vhi = v.min(0), v.max(0)
    lo = (vlo - np.float32(margin)).astype(np.float32)
    dims = tuple(int(np.ceil((float(vhi[a]) - float(vlo[a]) + 2 * margin) / dx)) + 1 for a in range(3))
    return dims, lo


def generate_on(vertices, faces, cfg: SphBoundaryMapConfig, dims, lo):
    """As generate, on a caller-given grid (dims nodes from lo, spacing
    cfg.dx); the dict also holds the mesh's bounding box (mesh_lo,
    mesh_hi)."""
    m, keep = _mesh(vertices, faces)
    dims = tuple(int(x) for x in dims)
    lo = np.asarray(lo, np.float32)
    shape = (dims[2], dims[1], dims[0])
    d = np.empty(shape, np.float32)
    n = np.empty(shape + (3,), np.float32)
    p = np.empty(shape, np.uint8)
    ms = C.c_float()
    _check(lib.sph_boundary_map_generate_on(C.byref(m), C.byref(cfg), (C.c_int32 * 3)(*dims),
                                            (C.c_float * 3)(*(float(x) for x in lo)), as_fptr(d), as_fptr(n),
                                            p.ctypes.data_as(C.POINTER(C.c_uint8)), C.byref(ms)), "generate_on")
    v = keep[0]
    hi = (lo + np.float32(cfg.dx) * (np.array(dims, np.float32) - 1)).astype(np.float32)
    return {"dims": dims, "lo": lo, "hi": hi, "distance": d, "normal": n, "present": p, "ms": float(ms.value),
            "mesh_lo": v.min(0), "mesh_hi": v.max(0)}


def save(path: str, cfg, bm: dict) -> None:
    _check(lib.sph_boundary_map_save(path.encode(), C.byref(cfg), (C.c_int32 * 3)(*bm["dims"]),
                                     (C.c_float * 3)(*bm["lo"]), (C.c_float * 3)(*bm["hi"]),
                                     as_fptr(np.ascontiguousarray(bm["distance"], np.float32)),
                                     as_fptr(np.ascontiguousarray(bm["normal"], np.float32)),
                                     np.ascontiguousarray(bm["present"], np.uint8).ctypes.data_as(C.POINTER(C.c_uint8))),
           "save")


def load(path: str):
    cfg = SphBoundaryMapConfig()
    dims, lo, hi = (C.c_int32 * 3)(), (C.c_float * 3)(), (C.c_float * 3)()
    _check(lib.sph_boundary_map_load(path.encode(), C.byref(cfg), dims, lo, hi, None, None, None), "load header")
    shape = (dims[2], dims[1], dims[0])
    d = np.empty(shape, np.float32)
    n = np.empty(shape + (3,), np.float32)
    p = np.empty(shape, np.uint8)
    _check(lib.sph_boundary_map_load(path.encode(), C.byref(cfg), dims, lo, hi, as_fptr(d), as_fptr(n),
                                     p.ctypes.data_as(C.POINTER(C.c_uint8))), "load")
    return cfg, {"dims": tuple(dims), "lo": np.array(lo[:], np.float32), "hi": np.array(hi[:], np.float32),
                 "distance": d, "normal": n, "present": p}
