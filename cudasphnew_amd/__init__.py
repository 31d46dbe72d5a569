"""MI355X-native SPH time step of CUDASPHNEW.

The product is the HIP library libsph_hip.so (C ABI in include/sph.h).  This
package holds its build recipe, ctypes bindings and a Python mirror of the
reference's ParticleSimulation interface.
"""
from ._native import (Config, Field, SceneDesc, SolidKind, SphBoundary, SphCounts, SphError, SphParams,
                      SphScene, Status, lib)
from .scenes import (Scene, boundary_of, config_desc, example01_boundary, make_scene, params_of,
                     particles_of, scene_from_desc, scene_sdf)
from .simulation import Batch, ParticleSimulation
from .neural import NeuralSDF, boundary_box, sdf_samples

__all__ = [
    "Batch", "Config", "Field", "NeuralSDF", "ParticleSimulation", "Scene", "SceneDesc", "SolidKind", "SphBoundary",
    "SphCounts", "SphError", "SphParams", "SphScene", "Status", "boundary_of", "config_desc",
    "example01_boundary", "lib", "make_scene", "params_of", "particles_of", "scene_from_desc",
    "scene_sdf", "boundary_box", "sdf_samples",
]


def device_count() -> int:
    return int(lib.sph_device_count())
