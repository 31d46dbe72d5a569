"""Python mirror of the reference's host interface (class ParticleSimulation,
particle_simulation.h:113-324), calling the HIP library through the C ABI.

Method names follow the reference (Init, Bind, Advance, AdvanceTwoScale,
GetParticleState, GetNumParticles, ...); errors raise SphError, the way the
reference throws std::runtime_error from CUDA_SAFE_CALL (util.h:24-39).
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from ._native import (Field, SphCounts, SphParams, SphRenderView, SphTuning, SPH_MAX_STAGES, as_fptr, check,
                      lib)
from .scenes import Scene, make_scene

_FLOAT3 = {Field.POSITION, Field.VELOCITY, Field.ACCEL, Field.SUB_POSITION, Field.SUB_VELOCITY,
           Field.SUB_ACCEL, Field.SUB_INPUT_POSITION, Field.SUB_INPUT_VELOCITY}
_FLOAT1 = {Field.DENSITY, Field.PRESSURE, Field.SUB_DENSITY, Field.SUB_PRESSURE}
_INT = {Field.SORTED_KEYS, Field.SORTED_IDS, Field.CELL_START, Field.CELL_END, Field.SUB_IDS,
        Field.SPLIT_IDS, Field.BOUNDARY_IDS, Field.DEFAULT_IDS, Field.SUB_SORTED_KEYS,
        Field.SUB_SORTED_IDS}


def _field_array(field: Field, n: int, ncells: int, counts: SphCounts, nsubcells: int = 0):
    nsub = counts.num_sub
    if field in (Field.CELL_START, Field.CELL_END):
        return np.empty(ncells, np.int32)
    if field in (Field.SUB_CELL_START, Field.SUB_CELL_END):
        return np.empty(nsubcells, np.int32)
    if field in (Field.STATE, Field.WALL_SOURCE):
        return np.empty(n, np.uint8)
    count = {Field.SUB_POSITION: nsub, Field.SUB_VELOCITY: nsub, Field.SUB_ACCEL: nsub,
             Field.SUB_INPUT_POSITION: nsub, Field.SUB_INPUT_VELOCITY: nsub,
             Field.SUB_DENSITY: nsub, Field.SUB_PRESSURE: nsub, Field.SUB_IDS: nsub,
             Field.SUB_SORTED_KEYS: nsub, Field.SUB_SORTED_IDS: nsub,
             Field.SPLIT_IDS: counts.num_split, Field.BOUNDARY_IDS: counts.num_boundary,
             Field.DEFAULT_IDS: counts.num_default}.get(field, n)
    if field in _FLOAT3:
        return np.empty((count, 3), np.float32)
    if field in _FLOAT1:
        return np.empty(count, np.float32)
    return np.empty(count, np.int32)


class ParticleSimulation:
    """One simulation handle (sph_handle)."""

    def __init__(self, scene: Scene):
        self.scene = scene
        self._h = None

    # -- construction (Example01 + Init + Bind, cu:1745,1882,2078) --------
    @classmethod
    def Example01(cls) -> "ParticleSimulation":
        return cls(make_scene("example01"))

    @classmethod
    def from_config(cls, config, index: int = 0, seed: int = 1) -> "ParticleSimulation":
        return cls(make_scene(config, index, seed))

    def Init(self) -> "ParticleSimulation":
        if self._h is not None:
            lib.sph_destroy(self._h)
            self._h = None
        h = C.c_void_p()
        sc = self.scene.sph_scene()
        st = lib.sph_create(C.byref(self.scene.params), C.byref(sc), C.byref(h))
        check(st, None)
        self._h = h
        return self

    def Bind(self) -> None:
        """Parameters are bound to the handle at Init / set_params; kept for
        interface parity with ParticleSimulation::Bind (cu:2078-2083)."""
        self._need()

    def close(self) -> None:
        if self._h is not None:
            lib.sph_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        if self._h is None:
            self.Init()
        return self

    def __exit__(self, *exc):
        self.close()

    def _need(self):
        if self._h is None:
            raise RuntimeError("call Init() first")
        return self._h

    # -- stepping ----------------------------------------------------------
    def Advance(self, n: int = 1) -> None:
        check(lib.sph_step(self._need(), int(n)), self._h)

    def AdvanceTwoScale(self, n: int = 1) -> None:
        check(lib.sph_step_two_scale(self._need(), int(n)), self._h)

    def AdvanceSubParticles(self, n: int = 1) -> None:
        """AdvanceSubParticles (cu:2119-2143): the sub-particles of the last
        two-scale step advance alone."""
        check(lib.sph_step_sub_particles(self._need(), int(n)), self._h)

    def sync(self) -> None:
        check(lib.sph_sync(self._need()), self._h)

    # -- parameters ----------------------------------------------------------
    def get_params(self) -> SphParams:
        p = SphParams()
        check(lib.sph_get_params(self._need(), C.byref(p)), self._h)
        return p

    def set_params(self, p: SphParams) -> None:
        check(lib.sph_set_params(self._need(), C.byref(p)), self._h)

    def SetNPartThresh(self, dval: float) -> None:
        p = self.get_params()
        p.n_part_thresh += dval
        self.set_params(p)

    # -- read back -----------------------------------------------------------
    def counts(self) -> SphCounts:
        c = SphCounts()
        check(lib.sph_get_counts(self._need(), C.byref(c)), self._h)
        return c

    def num_cells(self) -> int:
        return int(lib.sph_num_cells(self._need()))

    def read(self, field) -> np.ndarray:
        field = Field(field)
        nsubcells = int(np.prod([self.scene.params.grid_dim_sub[a] for a in range(3)]))
        out = _field_array(field, self.scene.n, self.num_cells(), self.counts(), nsubcells)
        check(lib.sph_read_field(self._need(), int(field), out.ctypes.data_as(C.c_void_p), out.nbytes),
              self._h)
        return out

    def positions(self) -> np.ndarray:
        out = np.empty((self.scene.n, 3), np.float32)
        check(lib.sph_read_positions(self._need(), as_fptr(out), self.scene.n), self._h)
        return out

    def map_positions(self):
        """(device pointer to cell-sorted float4 positions, device pointer to
        the original id of each slot)."""
        p, o = C.c_void_p(), C.c_void_p()
        check(lib.sph_map_positions(self._need(), C.byref(p), C.byref(o)), self._h)
        return p.value, o.value

    def map(self) -> dict:
        """ParticleSimulation::map (cu:2685-2700): device pointers and counts
        of the render buffers -- positions by original id, sub-particle
        positions by sub id (float4 x, y, z, density), the default-state id
        list and the active sub-particle id list (regular first)."""
        v = SphRenderView()
        check(lib.sph_map_render(self._need(), C.byref(v)), self._h)
        return {f: getattr(v, f) for f, _ in SphRenderView._fields_}

    def extract_surface(self):
        """extractSurfaceParticles (declared at particle_simulation.h:224,
        defined here; see sph_extract_surface): (flags by original id,
        uint8 0/1; number flagged)."""
        flags = np.zeros(self.scene.n, dtype=np.uint8)
        cnt = C.c_uint32()
        check(lib.sph_extract_surface(self._need(), flags.ctypes.data, C.byref(cnt)), self._h)
        return flags, int(cnt.value)

    def unmap(self) -> None:
        check(lib.sph_unmap_render(self._need()), self._h)

    def GetParticleState(self) -> np.ndarray:
        return self.read(Field.STATE)

    def GetNumParticles(self) -> int:
        return int(self.counts().num_particles)

    def GetNumTimesSteps(self) -> int:
        return int(self.counts().num_steps)

    def GetNumParticlesDefault(self) -> int:
        return int(self.counts().num_default)

    def GetNumSubParticles(self) -> int:
        return int(self.counts().num_sub)

    def GetNumSubParticlesRegular(self) -> int:
        return 8 * int(self.counts().num_split)

    def GetNumSubParticlesBoundary(self) -> int:
        return 8 * int(self.counts().num_boundary)

    def GetParticleRadius(self) -> float:
        # (3 V / (4 pi N))^(1/3) (cu:2227-2231)
        p = self.scene.params
        return float((3.0 * p.fluid_volume / (4.0 * np.pi * p.num_particles)) ** (1.0 / 3.0))

    def GetSubParticleRadius(self) -> float:
        return 0.5 * self.GetParticleRadius()

    def GetSizeMemoryGPU(self) -> int:
        return int(lib.sph_device_bytes(self._need()))

    # -- debug dumps (SaveInfoTable / SaveParticleInfo, cu:2709-2852) -------
    def SaveInfoTable(self, filename: str) -> None:
        """ParticleSimulation::SaveInfoTable (cu:2709-2780), same format."""
        check(lib.sph_save_info_table(self._need(), os.fsencode(filename)), self._h)

    def SaveParticleInfo(self, filename: str) -> None:
        """ParticleSimulation::SaveParticleInfo (cu:2782-2852), same format."""
        check(lib.sph_save_particle_info(self._need(), os.fsencode(filename)), self._h)

    # -- timing ----------------------------------------------------------------
    def enable_timing(self, on: bool = True) -> None:
        check(lib.sph_enable_timing(self._need(), 1 if on else 0), self._h)

    def stage_times(self) -> dict:
        ms = (C.c_float * SPH_MAX_STAGES)()
        cnt = C.c_int32()
        check(lib.sph_stage_times(self._need(), ms, C.byref(cnt)), self._h)
        return {lib.sph_stage_name(i).decode(): float(ms[i]) for i in range(cnt.value)}

    def use_graph(self, enable: bool = True) -> None:
        """Replay the steady-state plain step from a captured hipGraph."""
        check(lib.sph_use_graph(self._need(), 1 if enable else 0), self._h)

    def time_steps(self, n: int, two_scale: bool = False) -> float:
        ms = C.c_float()
        check(lib.sph_time_steps(self._need(), int(n), 1 if two_scale else 0, C.byref(ms)), self._h)
        return float(ms.value)


    # -- learned boundary -------------------------------------------------------
    def attach_neural(self, net, val=None, cfg=None) -> dict:
        from .neural import attach
        return attach(self, net, self.scene.desc, 0, val, cfg)

    def detach_neural(self) -> None:
        check(lib.sph_detach_neural_boundary(self._need()), self._h)
        self._neural = None

    # -- kernel tuning (variants of the density / force passes) ---------------
    def set_tuning(self, density_lanes: int = 1, force_lanes: int = 1, hit_masks: int = 0,
                   hit_list: int = 48, block: int = 256) -> None:
        t = SphTuning(int(density_lanes), int(force_lanes), int(hit_masks), int(hit_list), int(block))
        check(lib.sph_set_tuning(self._need(), C.byref(t)), self._h)

    def get_tuning(self) -> dict:
        t = SphTuning()
        check(lib.sph_get_tuning(self._need(), C.byref(t)), self._h)
        return _tuning_dict(t)

    def autotune(self, reps: int = 5) -> dict:
        t = SphTuning()
        check(lib.sph_autotune(self._need(), int(reps), C.byref(t)), self._h)
        return _tuning_dict(t)

    def set_autotune_interval(self, steps: int) -> None:
        """Re-run the autotuner before every `steps`-th step (0: off)."""
        check(lib.sph_set_autotune_interval(self._need(), int(steps)), self._h)


def _tuning_dict(t: SphTuning) -> dict:
    return {f: int(getattr(t, f)) for f, _ in SphTuning._fields_}


class Batch:
    """Many independent scenes advanced together (sph_batch)."""

    def __init__(self, scenes):
        self.scenes = list(scenes)
        n = len(self.scenes)
        self._params = (SphParams * n)(*[s.params for s in self.scenes])
        from ._native import SphScene
        self._sc = (SphScene * n)(*[s.sph_scene() for s in self.scenes])
        b = C.c_void_p()
        st = lib.sph_batch_create(n, self._params, self._sc, C.byref(b))
        check(st, None)
        self._b = b

    def close(self):
        if getattr(self, "_b", None) is not None:
            lib.sph_batch_destroy(self._b)
            self._b = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def step(self, n: int = 1) -> None:
        check(lib.sph_batch_step(self._b, int(n)), batch=self._b)

    def sync(self) -> None:
        check(lib.sph_batch_sync(self._b), batch=self._b)

    def use_graph(self, on: bool = True) -> None:
        check(lib.sph_batch_use_graph(self._b, 1 if on else 0), batch=self._b)

    def num_particles(self) -> int:
        return int(lib.sph_batch_num_particles(self._b))

    def read(self, scene: int, field) -> np.ndarray:
        field = Field(field)
        sc = self.scenes[scene]
        ncells = int(np.prod([sc.params.grid_dim[a] for a in range(3)]))
        out = _field_array(field, sc.n, ncells, SphCounts(num_particles=sc.n))
        check(lib.sph_batch_read_field(self._b, int(scene), int(field), out.ctypes.data_as(C.c_void_p),
                                       out.nbytes), batch=self._b)
        return out

    def time_steps(self, n: int) -> float:
        ms = C.c_float()
        check(lib.sph_batch_time_steps(self._b, int(n), C.byref(ms)), batch=self._b)
        return float(ms.value)

    def enable_timing(self, on: bool = True) -> None:
        check(lib.sph_batch_enable_timing(self._b, 1 if on else 0), batch=self._b)

    def stage_times(self) -> dict:
        ms = (C.c_float * SPH_MAX_STAGES)()
        cnt = C.c_int32()
        check(lib.sph_batch_stage_times(self._b, ms, C.byref(cnt)), batch=self._b)
        return {lib.sph_stage_name(i).decode(): float(ms[i]) for i in range(cnt.value)}

    def attach_neural(self, scene: int, net, val=None, cfg=None) -> dict:
        from .neural import attach
        return attach(self, net, self.scenes[scene].desc, scene, val, cfg)

    def set_tuning(self, density_lanes: int = 1, force_lanes: int = 1, hit_masks: int = 0,
                   hit_list: int = 48, block: int = 256) -> None:
        t = SphTuning(int(density_lanes), int(force_lanes), int(hit_masks), int(hit_list), int(block))
        check(lib.sph_batch_set_tuning(self._b, C.byref(t)), batch=self._b)

    def get_tuning(self) -> dict:
        t = SphTuning()
        check(lib.sph_batch_get_tuning(self._b, C.byref(t)), batch=self._b)
        return _tuning_dict(t)

    def autotune(self, reps: int = 5) -> dict:
        t = SphTuning()
        check(lib.sph_batch_autotune(self._b, int(reps), C.byref(t)), batch=self._b)
        return _tuning_dict(t)

    def set_autotune_interval(self, steps: int) -> None:
        """Re-run the autotuner before every `steps`-th step (0: off)."""
        check(lib.sph_batch_set_autotune_interval(self._b, int(steps)), batch=self._b)
