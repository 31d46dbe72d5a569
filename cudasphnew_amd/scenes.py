"""Scene construction: named workloads, particles and boundary grids.

Thin wrapper over the host-side C functions sph_scene_* (scene.cpp), which
replace Example01 / create_particle_box / set_simulation_domain and the
SphInComplexShapes grid builder of the reference.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from ._native import (Config, SceneDesc, SphBoundary, SphParams, SphScene, as_fptr, check, lib)


@dataclass
class Scene:
    desc: SceneDesc
    params: SphParams
    positions: np.ndarray              # (N, 3) float32, original-id order
    distance: np.ndarray               # (nz, ny, nx) float32
    density: np.ndarray
    viscosity: np.ndarray
    velocities: np.ndarray | None = None
    _keep: list = field(default_factory=list, repr=False)

    @property
    def n(self) -> int:
        return int(self.positions.shape[0])

    @property
    def boundary_dims(self):
        return tuple(int(v) for v in self.desc.boundary_dims)

    def sph_boundary(self) -> SphBoundary:
        b = SphBoundary()
        for a in range(3):
            b.origin[a] = self.desc.boundary_origin[a]
            b.dims[a] = self.desc.boundary_dims[a]
        b.spacing = self.desc.boundary_spacing
        b.rest_distance = self.desc.rest_distance
        for name in ("distance", "density", "viscosity"):
            arr = np.ascontiguousarray(getattr(self, name), dtype=np.float32)
            setattr(self, name, arr)
            setattr(b, name, as_fptr(arr))
        return b

    def sph_scene(self) -> SphScene:
        s = SphScene()
        pos = np.ascontiguousarray(self.positions, dtype=np.float32)
        self.positions = pos
        s.num_particles = pos.shape[0]
        s.positions = as_fptr(pos)
        if self.velocities is not None:
            vel = np.ascontiguousarray(self.velocities, dtype=np.float32)
            self.velocities = vel
            s.velocities = as_fptr(vel)
        s.boundary = self.sph_boundary()
        self._keep = [pos]
        return s

    def with_positions(self, positions: np.ndarray, velocities: np.ndarray | None = None) -> "Scene":
        """Same boundary and physics, other particles (params follow N)."""
        p = SphParams.from_buffer_copy(self.params)
        p.num_particles = positions.shape[0]
        return Scene(self.desc, p, np.ascontiguousarray(positions, np.float32), self.distance,
                     self.density, self.viscosity,
                     None if velocities is None else np.ascontiguousarray(velocities, np.float32))


def config_desc(config, index: int = 0, seed: int = 1) -> SceneDesc:
    if isinstance(config, str):
        config = Config[config.upper()]
    d = SceneDesc()
    check(lib.sph_scene_config(int(config), int(index), int(seed), C.byref(d)))
    return d


def params_of(desc: SceneDesc) -> SphParams:
    p = SphParams()
    check(lib.sph_scene_params(C.byref(desc), C.byref(p)))
    return p


def particles_of(desc: SceneDesc) -> np.ndarray:
    n = C.c_uint32()
    check(lib.sph_scene_particles(C.byref(desc), None, 0, C.byref(n)))
    pos = np.empty((n.value, 3), np.float32)
    check(lib.sph_scene_particles(C.byref(desc), as_fptr(pos), n.value, C.byref(n)))
    return pos


def boundary_of(desc: SceneDesc):
    nx, ny, nz = (int(v) for v in desc.boundary_dims)
    out = [np.empty((nz, ny, nx), np.float32) for _ in range(3)]
    check(lib.sph_scene_boundary(C.byref(desc), *(as_fptr(a) for a in out)))
    return out


def scene_from_desc(desc: SceneDesc) -> Scene:
    params = params_of(desc)
    pos = particles_of(desc)
    d, w, v = boundary_of(desc)
    return Scene(desc, params, pos, d, w, v)


def make_scene(config, index: int = 0, seed: int = 1) -> Scene:
    """Named workload: 'dam16k' (16,384 particles), 'dam1m' (1,048,576),
    'two_scale' (262,144 coarse), 'batch' (scene `index` of the 64-scene
    batch), 'example01' (the reference's Example01)."""
    return scene_from_desc(config_desc(config, index, seed))


def scene_sdf(desc: SceneDesc, xyz: np.ndarray) -> np.ndarray:
    pts = np.ascontiguousarray(xyz, np.float32).reshape(-1, 3)
    out = np.empty(pts.shape[0], np.float32)
    check(lib.sph_scene_sdf(C.byref(desc), as_fptr(pts), pts.shape[0], as_fptr(out)))
    return out


def example01_boundary():
    """Distance/density/viscosity grids of Example01's SetBox call
    (particle_simulation.cu:2592-2612), shape (nz, ny, nx)."""
    dims = (C.c_int32 * 3)()
    origin = (C.c_float * 3)()
    sp = C.c_float()
    r0 = C.c_float()
    check(lib.sph_example01_boundary(dims, origin, C.byref(sp), C.byref(r0), None, None, None))
    nx, ny, nz = dims
    out = [np.empty((nz, ny, nx), np.float32) for _ in range(3)]
    check(lib.sph_example01_boundary(dims, origin, C.byref(sp), C.byref(r0), *(as_fptr(a) for a in out)))
    return {"dims": (nx, ny, nz), "origin": tuple(origin), "spacing": sp.value, "rest_distance": r0.value,
            "distance": out[0], "density": out[1], "viscosity": out[2]}
