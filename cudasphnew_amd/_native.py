"""ctypes bindings of the C ABI declared in include/sph.h.

The library is loaded from the package directory (libsph_hip.so, built by
``python -m cudasphnew_amd.build``).  There is no fallback: if the library is
missing, importing this module raises.
"""
from __future__ import annotations

import ctypes as C
import enum
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libsph_hip.so")

f3 = C.c_float * 3
i3 = C.c_int32 * 3
u3 = C.c_uint32 * 3
fptr = C.POINTER(C.c_float)

SPH_EMPTY_CELL = 0xFFFFFFFF
SPH_MAX_STAGES = 16
SPH_MAX_OBSTACLES = 8


class Status(enum.IntEnum):
    OK = 0
    INVALID_ARG = 1
    OUT_OF_MEMORY = 2
    DEVICE_ERROR = 3
    NONFINITE_STATE = 4
    OUT_OF_DOMAIN = 5
    NO_DEVICE = 6
    CAPACITY_EXCEEDED = 7


class Field(enum.IntEnum):
    POSITION = 0
    VELOCITY = 1
    DENSITY = 2
    PRESSURE = 3
    ACCEL = 4
    STATE = 5
    SORTED_KEYS = 6
    SORTED_IDS = 7
    CELL_START = 8
    CELL_END = 9
    SUB_POSITION = 10
    SUB_VELOCITY = 11
    SUB_DENSITY = 12
    SUB_PRESSURE = 13
    SUB_ACCEL = 14
    SUB_IDS = 15
    SPLIT_IDS = 16
    BOUNDARY_IDS = 17
    DEFAULT_IDS = 18
    SUB_SORTED_KEYS = 19
    SUB_SORTED_IDS = 20
    WALL_SOURCE = 21
    SUB_CELL_START = 22
    SUB_CELL_END = 23
    SUB_INPUT_POSITION = 24
    SUB_INPUT_VELOCITY = 25


class Config(enum.IntEnum):
    EXAMPLE01 = 0
    DAM16K = 1
    DAM1M = 2
    TWO_SCALE = 3
    BATCH = 5


class SolidKind(enum.IntEnum):
    NONE = 0
    SPHERE = 1
    CONE = 2
    BOX = 3


class SphParams(C.Structure):
    _fields_ = [
        ("grid_origin", f3), ("grid_spacing", C.c_float), ("grid_dim", i3),
        ("grid_spacing_sub", C.c_float), ("grid_dim_sub", i3),
        ("num_particles", C.c_uint32), ("compact_support", C.c_float),
        ("compact_support_sub", C.c_float), ("kernel_particles", C.c_int32),
        ("time_step", C.c_float), ("time_step_sub", C.c_float),
        ("particle_mass", C.c_float), ("sub_particle_mass", C.c_float),
        ("gas_stiffness", C.c_float), ("rest_density", C.c_float),
        ("dynamic_viscosity", C.c_float), ("gravity", C.c_float),
        ("tension_coefficient", C.c_float), ("norm_thresh", C.c_float),
        ("fluid_volume", C.c_float),
        ("poly6", C.c_float), ("grad_poly6", C.c_float), ("lapl_poly6", C.c_float),
        ("grad_spiky", C.c_float), ("lapl_visc", C.c_float),
        ("poly6_sub", C.c_float), ("grad_poly6_sub", C.c_float),
        ("lapl_poly6_sub", C.c_float), ("grad_spiky_sub", C.c_float),
        ("lapl_visc_sub", C.c_float),
        ("cm_distance_thresh", C.c_float), ("n_part_thresh", C.c_float),
        ("box_dim", f3), ("box_cen", f3), ("restitution", C.c_float),
        ("collision", C.c_int32), ("split_lo", f3), ("split_hi", f3),
        ("sub_placement", C.c_int32), ("max_sub_particles", C.c_uint32),
    ]


class SphBoundary(C.Structure):
    _fields_ = [
        ("origin", f3), ("spacing", C.c_float), ("dims", i3),
        ("rest_distance", C.c_float), ("distance", fptr), ("density", fptr),
        ("viscosity", fptr),
    ]


class SphDeviceBoundary(C.Structure):
    """sph_device_boundary: the boundary grids in device memory (grid =
    float4 per node, distance = float per node); pointers as integers."""
    _fields_ = [
        ("origin", f3), ("spacing", C.c_float), ("dims", i3),
        ("rest_distance", C.c_float), ("grid", C.c_void_p), ("distance", C.c_void_p),
    ]


class SphScene(C.Structure):
    _fields_ = [
        ("num_particles", C.c_uint32), ("positions", fptr), ("velocities", fptr),
        ("boundary", SphBoundary),
    ]


class SphRenderView(C.Structure):
    _fields_ = [
        ("positions", C.c_void_p), ("sub_positions", C.c_void_p),
        ("default_ids", C.c_void_p), ("sub_ids", C.c_void_p),
        ("num_particles", C.c_uint32), ("num_default", C.c_uint32),
        ("num_sub_regular", C.c_uint32), ("num_sub_boundary", C.c_uint32),
    ]


class SphCounts(C.Structure):
    _fields_ = [
        ("num_particles", C.c_uint32), ("num_split", C.c_uint32),
        ("num_boundary", C.c_uint32), ("num_default", C.c_uint32),
        ("num_sub", C.c_uint32), ("num_steps", C.c_uint32),
    ]


class SphTuning(C.Structure):
    _fields_ = [("density_lanes", C.c_int32), ("force_lanes", C.c_int32),
                ("hit_masks", C.c_int32), ("hit_list", C.c_int32), ("block", C.c_int32)]


class SphSolid(C.Structure):
    _fields_ = [("kind", C.c_int32), ("center", f3), ("a", f3)]


class SceneDesc(C.Structure):
    _fields_ = [
        ("lattice", u3), ("lattice_origin", f3), ("lattice_spacing", C.c_float),
        ("jitter", C.c_float), ("seed", C.c_uint64),
        ("box_lo", f3), ("box_hi", f3), ("num_obstacles", C.c_int32),
        ("obstacles", SphSolid * SPH_MAX_OBSTACLES),
        ("h", C.c_float), ("rest_density", C.c_float), ("particle_mass", C.c_float),
        ("gas_stiffness", C.c_float), ("viscosity", C.c_float), ("gravity", C.c_float),
        ("tension", C.c_float), ("norm_thresh", C.c_float), ("dt", C.c_float),
        ("dt_sub", C.c_float), ("cell_lo", f3), ("cell_hi", f3),
        ("boundary_origin", f3), ("boundary_spacing", C.c_float), ("boundary_dims", i3),
        ("rest_distance", C.c_float), ("boundary_support", C.c_float),
        ("ghost_spacing", C.c_float), ("collision", C.c_int32), ("restitution", C.c_float),
        ("split_lo", f3), ("split_hi", f3), ("sub_placement", C.c_int32),
    ]


SPH_MLP_NUM_PARAMS = 16641


class SphMlpConfig(C.Structure):
    _fields_ = [
        ("box_lo", f3), ("box_hi", f3), ("out_scale", C.c_float), ("band", C.c_float), ("seed", C.c_uint64),
        ("lr", C.c_float), ("beta1", C.c_float), ("beta2", C.c_float), ("eps", C.c_float),
        ("lr_min", C.c_float), ("decay_steps", C.c_int32),
    ]


class SphNeuralConfig(C.Structure):
    _fields_ = [("tolerance", C.c_float), ("band", C.c_float), ("normal_tolerance", C.c_float),
                ("brick", C.c_int32)]


class SphNeuralReport(C.Structure):
    _fields_ = [
        ("cells_none", C.c_uint64), ("cells_neural", C.c_uint64), ("cells_grid", C.c_uint64),
        ("bricks_checked", C.c_uint32), ("bricks_fallback", C.c_uint32),
        ("max_error", C.c_float), ("mean_error", C.c_float), ("samples_in_band", C.c_uint64),
        ("mean_normal_error", C.c_float), ("bricks_fallback_normal", C.c_uint32),
    ]


class SphMesh(C.Structure):
    _fields_ = [("vertices", C.POINTER(C.c_float)), ("num_vertices", C.c_uint32),
                ("faces", C.POINTER(C.c_uint32)), ("num_faces", C.c_uint32)]


class SphBoundaryMapConfig(C.Structure):
    _fields_ = [("rest_distance", C.c_float), ("compact_support", C.c_float), ("dx", C.c_float)]


class SphMapSolid(C.Structure):
    _fields_ = [("dims", i3), ("lo", f3), ("dx", C.c_float), ("band", C.c_float),
                ("distance", C.POINTER(C.c_float)), ("present", C.POINTER(C.c_uint8)),
                ("solid_lo", f3), ("solid_hi", f3)]


class SphError(RuntimeError):
    """Raised for a non-OK sph_status (the reference throws
    std::runtime_error from CUDA_SAFE_CALL, util.h:24-39)."""

    def __init__(self, status: int, message: str = ""):
        self.status = Status(status)
        super().__init__(f"{self.status.name}: {message}")


def _load():
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is missing; build it with `python cudasphnew_amd/build.py`")
    return C.CDLL(LIB_PATH)


lib = _load()

_vp = C.c_void_p
_st = C.c_int
# (name, restype, argtypes) of every exported symbol of include/sph.h
SIGNATURES = [
    ("sph_create", _st, [C.POINTER(SphParams), C.POINTER(SphScene), C.POINTER(_vp)]),
    ("sph_destroy", _st, [_vp]),
    ("sph_step", _st, [_vp, C.c_int32]),
    ("sph_step_two_scale", _st, [_vp, C.c_int32]),
    ("sph_step_sub_particles", _st, [_vp, C.c_int32]),
    ("sph_sync", _st, [_vp]),
    ("sph_set_params", _st, [_vp, C.POINTER(SphParams)]),
    ("sph_get_params", _st, [_vp, C.POINTER(SphParams)]),
    ("sph_read_positions", _st, [_vp, fptr, C.c_size_t]),
    ("sph_read_field", _st, [_vp, C.c_int, _vp, C.c_size_t]),
    ("sph_map_positions", _st, [_vp, C.POINTER(_vp), C.POINTER(_vp)]),
    ("sph_map_render", _st, [_vp, C.POINTER(SphRenderView)]),
    ("sph_use_graph", _st, [_vp, C.c_int32]),
    ("sph_unmap_render", _st, [_vp]),
    ("sph_get_counts", _st, [_vp, C.POINTER(SphCounts)]),
    ("sph_save_info_table", _st, [_vp, C.c_char_p]),
    ("sph_extract_surface", _st, [_vp, C.c_void_p, C.POINTER(C.c_uint32)]),
    ("sph_save_particle_info", _st, [_vp, C.c_char_p]),
    ("sph_num_cells", C.c_size_t, [_vp]),
    ("sph_device_bytes", C.c_size_t, [_vp]),
    ("sph_last_error_message", C.c_char_p, [_vp]),
    ("sph_enable_timing", _st, [_vp, C.c_int32]),
    ("sph_stage_times", _st, [_vp, fptr, C.POINTER(C.c_int32)]),
    ("sph_stage_name", C.c_char_p, [C.c_int32]),
    ("sph_time_steps", _st, [_vp, C.c_int32, C.c_int32, fptr]),
    ("sph_set_autotune_interval", _st, [_vp, C.c_int32]),
    ("sph_set_tuning", _st, [_vp, C.POINTER(SphTuning)]),
    ("sph_get_tuning", _st, [_vp, C.POINTER(SphTuning)]),
    ("sph_autotune", _st, [_vp, C.c_int32, C.POINTER(SphTuning)]),
    ("sph_batch_create", _st, [C.c_int32, C.POINTER(SphParams), C.POINTER(SphScene), C.POINTER(_vp)]),
    ("sph_batch_destroy", _st, [_vp]),
    ("sph_batch_step", _st, [_vp, C.c_int32]),
    ("sph_batch_sync", _st, [_vp]),
    ("sph_batch_read_field", _st, [_vp, C.c_int32, C.c_int, _vp, C.c_size_t]),
    ("sph_batch_time_steps", _st, [_vp, C.c_int32, fptr]),
    ("sph_batch_enable_timing", _st, [_vp, C.c_int32]),
    ("sph_batch_stage_times", _st, [_vp, fptr, C.POINTER(C.c_int32)]),
    ("sph_batch_use_graph", _st, [_vp, C.c_int32]),
    ("sph_batch_set_tuning", _st, [_vp, C.POINTER(SphTuning)]),
    ("sph_batch_get_tuning", _st, [_vp, C.POINTER(SphTuning)]),
    ("sph_batch_autotune", _st, [_vp, C.c_int32, C.POINTER(SphTuning)]),
    ("sph_batch_set_autotune_interval", _st, [_vp, C.c_int32]),
    ("sph_batch_num_particles", C.c_uint64, [_vp]),
    ("sph_batch_last_error_message", C.c_char_p, [_vp]),
    ("sph_scene_config", _st, [C.c_int32, C.c_int32, C.c_uint64, C.POINTER(SceneDesc)]),
    ("sph_scene_params", _st, [C.POINTER(SceneDesc), C.POINTER(SphParams)]),
    ("sph_scene_particles", _st, [C.POINTER(SceneDesc), fptr, C.c_uint32, C.POINTER(C.c_uint32)]),
    ("sph_scene_boundary", _st, [C.POINTER(SceneDesc), fptr, fptr, fptr]),
    ("sph_scene_sdf", _st, [C.POINTER(SceneDesc), fptr, C.c_size_t, fptr]),
    ("sph_scene_boundary_maps", _st, [C.POINTER(SceneDesc), C.POINTER(SphMapSolid), C.c_int32, fptr, fptr, fptr]),
    ("sph_scene_sdf_maps", _st, [C.POINTER(SceneDesc), C.POINTER(SphMapSolid), C.c_int32, fptr, C.c_size_t, fptr]),
    ("sph_example01_boundary", _st, [i3, f3, fptr, fptr, fptr, fptr, fptr]),
    ("sph_scene_grid_samples", _st, [C.POINTER(SceneDesc), C.c_uint64, C.c_size_t, C.c_float, C.c_float, fptr]),
    ("sph_scene_sdf_samples", _st, [C.POINTER(SceneDesc), C.c_uint64, C.c_size_t, C.c_float, C.c_float, fptr]),
    ("sph_scene_sdf_samples_maps", _st, [C.POINTER(SceneDesc), C.POINTER(SphMapSolid), C.c_int32, C.c_uint64,
                                         C.c_size_t, C.c_float, C.c_float, fptr]),
    ("sph_scene_grid_samples_maps", _st, [C.POINTER(SceneDesc), C.POINTER(SphMapSolid), C.c_int32, C.c_uint64,
                                          C.c_size_t, C.c_float, C.c_float, fptr]),
    ("sph_mlp_default_config", _st, [f3, f3, C.POINTER(SphMlpConfig)]),
    ("sph_mlp_create", _st, [C.POINTER(SphMlpConfig), C.POINTER(_vp)]),
    ("sph_mlp_destroy", _st, [_vp]),
    ("sph_mlp_get_config", _st, [_vp, C.POINTER(SphMlpConfig)]),
    ("sph_mlp_set_params", _st, [_vp, fptr, C.c_size_t]),
    ("sph_mlp_get_params", _st, [_vp, fptr, C.c_size_t]),
    ("sph_mlp_set_samples", _st, [_vp, fptr, C.c_size_t]),
    ("sph_mlp_gradient", _st, [_vp, C.c_uint32, C.c_int32, fptr, fptr]),
    ("sph_mlp_train", _st, [_vp, C.c_int32, C.c_int32, fptr]),
    ("sph_mlp_time_train", _st, [_vp, C.c_int32, C.c_int32, fptr]),
    ("sph_mlp_eval", _st, [_vp, fptr, C.c_size_t, fptr]),
    ("sph_mlp_last_error_message", C.c_char_p, [_vp]),
    ("sph_attach_neural_boundary", _st, [_vp, _vp, fptr, C.c_size_t, C.POINTER(SphNeuralConfig),
                                         C.POINTER(SphNeuralReport)]),
    ("sph_detach_neural_boundary", _st, [_vp]),
    ("sph_batch_attach_neural_boundary", _st, [_vp, C.c_int32, _vp, fptr, C.c_size_t,
                                               C.POINTER(SphNeuralConfig), C.POINTER(SphNeuralReport)]),
    ("sph_obj_read", _st, [C.c_char_p, fptr, C.POINTER(C.c_uint32), C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]),
    ("sph_mesh_fit", _st, [fptr, C.c_uint32, f3, f3]),
    ("sph_boundary_map_domain", _st, [C.POINTER(SphMesh), C.c_float, i3, f3, f3]),
    ("sph_boundary_map_generate", _st, [C.POINTER(SphMesh), C.POINTER(SphBoundaryMapConfig), fptr, fptr,
                                        C.POINTER(C.c_uint8), fptr]),
    ("sph_boundary_map_generate_on", _st, [C.POINTER(SphMesh), C.POINTER(SphBoundaryMapConfig), i3, f3, fptr,
                                           fptr, C.POINTER(C.c_uint8), fptr]),
    ("sph_boundary_map_save", _st, [C.c_char_p, C.POINTER(SphBoundaryMapConfig), i3, f3, f3, fptr, fptr,
                                    C.POINTER(C.c_uint8)]),
    ("sph_boundary_map_load", _st, [C.c_char_p, C.POINTER(SphBoundaryMapConfig), i3, f3, f3, fptr, fptr,
                                    C.POINTER(C.c_uint8)]),
    ("sph_abi_version", C.c_int32, []),
    ("sph_device_count", C.c_int32, []),
    ("sph_step_noalloc", _st, [C.POINTER(SphParams), C.POINTER(SphDeviceBoundary), C.c_uint32, C.c_void_p,
                               C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p]),
]

for _name, _res, _args in SIGNATURES:
    _fn = getattr(lib, _name)
    _fn.restype = _res
    _fn.argtypes = _args


def check(status: int, handle=None, batch=None) -> None:
    if status == Status.OK:
        return
    if batch is not None:
        msg = lib.sph_batch_last_error_message(batch)
    else:
        msg = lib.sph_last_error_message(handle)
    raise SphError(status, (msg or b"").decode(errors="replace"))


def as_fptr(a):
    return a.ctypes.data_as(fptr)
