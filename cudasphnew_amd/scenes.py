"""Scene construction: named workloads, particles and boundary grids.

Thin wrapper over the host-side C functions sph_scene_* (scene.cpp), which
replace Example01 / create_particle_box / set_simulation_domain and the
SphInComplexShapes grid builder of the reference.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from ._native import (Config, SceneDesc, SphBoundary, SphMapSolid, SphParams, SphScene, as_fptr, check, lib)


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


def ghost_spacing_of(desc: SceneDesc) -> float:
    """Spacing of the grid builder's ghost lattice: explicit, else
    (m/rho0)^(1/3)."""
    if desc.ghost_spacing > 0:
        return float(desc.ghost_spacing)
    m = desc.particle_mass if desc.particle_mass > 0 else desc.rest_density * desc.lattice_spacing ** 3
    return float((m / desc.rest_density) ** (1.0 / 3.0))


def map_solid(bm: dict, band: float, solid_lo=None, solid_hi=None):
    """sph_map_solid over a boundary map dict (meshsdf.generate_on /
    generate / load): the solid's box defaults to the mesh's (mesh_lo,
    mesh_hi) or the map's domain.  Returns (struct, arrays to keep alive)."""
    d = np.ascontiguousarray(bm["distance"], np.float32)
    p = np.ascontiguousarray(bm["present"], np.uint8)
    m = SphMapSolid()
    for a in range(3):
        m.dims[a] = int(bm["dims"][a])
        m.lo[a] = float(bm["lo"][a])
    m.dx = float(bm["dx"]) if "dx" in bm else float((bm["hi"][0] - bm["lo"][0]) / (bm["dims"][0] - 1))
    m.band = float(band)
    m.distance = as_fptr(d)
    m.present = p.ctypes.data_as(C.POINTER(C.c_uint8))
    if solid_lo is None:
        solid_lo = bm["mesh_lo"] if "mesh_lo" in bm else bm["lo"]
    if solid_hi is None:
        solid_hi = bm["mesh_hi"] if "mesh_hi" in bm else bm["hi"]
    for a in range(3):
        m.solid_lo[a] = float(solid_lo[a])
        m.solid_hi[a] = float(solid_hi[a])
    return m, (d, p)


def mesh_obstacle(vertices, faces, desc: SceneDesc, dx: float | None = None):
    """Boundary map of a mesh (already placed, e.g. by meshsdf.fit) made to
    serve as an obstacle of the scene `desc`: generated on the GPU on a grid
    of spacing dx (default half the scene's boundary spacing) over the
    mesh's box grown by the band, band = support + ghost spacing + 2 dx.
    Returns the map dict with 'dx' and 'band' added."""
    from . import meshsdf
    if dx is None:
        dx = float(desc.boundary_spacing) / 2
    band = float(desc.boundary_support) + ghost_spacing_of(desc) + 2.0 * dx
    cfg = meshsdf.config(float(desc.rest_distance), band, dx)
    dims, lo = meshsdf.grown_domain(vertices, dx, band + dx)
    bm = meshsdf.generate_on(vertices, faces, cfg, dims, lo)
    bm["dx"] = float(cfg.dx)
    bm["band"] = float(np.float32(band))
    return bm


def _maps(desc, maps):
    maps = list(maps or [])
    structs, keep = [], []
    for bm in maps:
        m, k = map_solid(bm, bm["band"]) if isinstance(bm, dict) else (bm, ())
        structs.append(m)
        keep.append(k)
    arr = (SphMapSolid * max(1, len(structs)))(*structs)
    return arr, len(structs), keep


def boundary_of(desc: SceneDesc, maps=None):
    """Distance, density and viscosity grids (nz, ny, nx) of the scene;
    `maps`: extra obstacles given as boundary map dicts (mesh_obstacle) or
    SphMapSolid structs."""
    nx, ny, nz = (int(v) for v in desc.boundary_dims)
    out = [np.empty((nz, ny, nx), np.float32) for _ in range(3)]
    if not maps:
        check(lib.sph_scene_boundary(C.byref(desc), *(as_fptr(a) for a in out)))
    else:
        arr, n, keep = _maps(desc, maps)
        check(lib.sph_scene_boundary_maps(C.byref(desc), arr, n, *(as_fptr(a) for a in out)))
    return out


def scene_from_desc(desc: SceneDesc, maps=None) -> Scene:
    params = params_of(desc)
    pos = particles_of(desc)
    d, w, v = boundary_of(desc, maps)
    return Scene(desc, params, pos, d, w, v)


def make_scene(config, index: int = 0, seed: int = 1) -> Scene:
    """Named workload: 'dam16k' (16,384 particles), 'dam1m' (1,048,576),
    'two_scale' (262,144 coarse), 'batch' (scene `index` of the 64-scene
    batch), 'example01' (the reference's Example01)."""
    return scene_from_desc(config_desc(config, index, seed))


def scene_sdf(desc: SceneDesc, xyz: np.ndarray, maps=None) -> np.ndarray:
    pts = np.ascontiguousarray(xyz, np.float32).reshape(-1, 3)
    out = np.empty(pts.shape[0], np.float32)
    if not maps:
        check(lib.sph_scene_sdf(C.byref(desc), as_fptr(pts), pts.shape[0], as_fptr(out)))
    else:
        arr, n, keep = _maps(desc, maps)
        check(lib.sph_scene_sdf_maps(C.byref(desc), arr, n, as_fptr(pts), pts.shape[0], as_fptr(out)))
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
