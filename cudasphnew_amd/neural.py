"""Learned boundary SDF (sph_mlp_*): training samples, GPU training and
evaluation of the Fourier-feature MLP that stands in for the distance grid.

The network, its parameter layout and the training loss are defined in
include/sph.h; the kernels are in csrc/neural.hip.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._native import (SPH_MLP_NUM_PARAMS, SceneDesc, SphError, SphMlpConfig, SphNeuralConfig, SphNeuralReport,
                      Status, as_fptr, lib)

NUM_PARAMS = SPH_MLP_NUM_PARAMS


def _check(status: int, handle=None) -> None:
    if status != Status.OK:
        msg = lib.sph_mlp_last_error_message(handle)
        raise SphError(status, (msg or b"").decode(errors="replace"))


def boundary_box(desc: SceneDesc):
    """The boundary grid's box: origin .. origin + (dims - 1) * spacing."""
    lo = np.array([desc.boundary_origin[a] for a in range(3)], np.float32)
    ext = np.array([(desc.boundary_dims[a] - 1) * desc.boundary_spacing for a in range(3)], np.float32)
    return lo, (lo + ext).astype(np.float32)


def _samples(plain_fn, maps_fn, desc, n, near_frac, near_width, seed, maps):
    if near_width is None:
        near_width = 3.0 * float(desc.h)
    out = np.empty((int(n), 4), np.float32)
    if not maps:
        st = plain_fn(C.byref(desc), int(seed), int(n), float(near_frac), float(near_width), as_fptr(out))
    else:
        from .scenes import _maps
        arr, nm, keep = _maps(desc, maps)
        st = maps_fn(C.byref(desc), arr, nm, int(seed), int(n), float(near_frac), float(near_width), as_fptr(out))
    if st != Status.OK:
        raise SphError(st, "sample generation failed")
    return out


def sdf_samples(desc: SceneDesc, n: int, near_frac: float = 0.3, near_width: float | None = None,
                seed: int = 1, maps=None) -> np.ndarray:
    """(n, 4) float32 (x, y, z, sdf) samples of the scene's boundary: a
    fraction near_frac within near_width (default 3 h) of the surface, the
    rest uniform over the boundary grid's box.  `maps`: obstacles given as
    boundary maps (scenes.mesh_obstacle), as for scenes.boundary_of."""
    return _samples(lib.sph_scene_sdf_samples, lib.sph_scene_sdf_samples_maps, desc, n, near_frac, near_width,
                    seed, maps)


def grid_samples(desc: SceneDesc, n: int, near_frac: float = 0.3, near_width: float | None = None,
                 seed: int = 1, maps=None) -> np.ndarray:
    """The points of sdf_samples labelled with the boundary grid's distance
    texture (the field the step samples)."""
    return _samples(lib.sph_scene_grid_samples, lib.sph_scene_grid_samples_maps, desc, n, near_frac, near_width,
                    seed, maps)


class NeuralSDF:
    """sph_mlp handle: FP32 master parameters, Adam state and BF16 copies on
    the device."""

    def __init__(self, box_lo, box_hi, seed: int = 1, out_scale: float | None = None, lr: float | None = None,
                 band: float = 0.0, lr_min: float = 0.0, decay_steps: int = 0):
        cfg = SphMlpConfig()
        lo = (C.c_float * 3)(*[float(v) for v in box_lo])
        hi = (C.c_float * 3)(*[float(v) for v in box_hi])
        _check(lib.sph_mlp_default_config(lo, hi, C.byref(cfg)))
        cfg.seed = int(seed)
        if out_scale is not None:
            cfg.out_scale = float(out_scale)
        if lr is not None:
            cfg.lr = float(lr)
        cfg.band = float(band)
        cfg.lr_min = float(lr_min)
        cfg.decay_steps = int(decay_steps)
        h = C.c_void_p()
        _check(lib.sph_mlp_create(C.byref(cfg), C.byref(h)))
        self._h = h
        self.config = cfg

    @classmethod
    def for_scene(cls, desc: SceneDesc, **kw) -> "NeuralSDF":
        lo, hi = boundary_box(desc)
        return cls(lo, hi, **kw)

    def close(self) -> None:
        if getattr(self, "_h", None) is not None:
            lib.sph_mlp_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def handle(self):
        return self._h

    def get_params(self) -> np.ndarray:
        out = np.empty(NUM_PARAMS, np.float32)
        _check(lib.sph_mlp_get_params(self._h, as_fptr(out), NUM_PARAMS), self._h)
        return out

    def set_params(self, theta) -> None:
        t = np.ascontiguousarray(theta, np.float32)
        if t.size != NUM_PARAMS:
            raise ValueError(f"need {NUM_PARAMS} parameters")
        _check(lib.sph_mlp_set_params(self._h, as_fptr(t), NUM_PARAMS), self._h)

    def set_samples(self, xyzd) -> None:
        s = np.ascontiguousarray(xyzd, np.float32).reshape(-1, 4)
        self._samples = s
        _check(lib.sph_mlp_set_samples(self._h, as_fptr(s), s.shape[0]), self._h)

    def gradient(self, first: int, batch: int):
        g = np.empty(NUM_PARAMS, np.float32)
        loss = C.c_float()
        _check(lib.sph_mlp_gradient(self._h, int(first), int(batch), as_fptr(g), C.byref(loss)), self._h)
        return float(loss.value), g

    def train(self, steps: int, batch: int = 65536) -> np.ndarray:
        losses = np.empty(max(int(steps), 1), np.float32)
        _check(lib.sph_mlp_train(self._h, int(steps), int(batch), as_fptr(losses)), self._h)
        return losses[:int(steps)]

    def time_train(self, steps: int, batch: int = 65536) -> float:
        ms = C.c_float()
        _check(lib.sph_mlp_time_train(self._h, int(steps), int(batch), C.byref(ms)), self._h)
        return float(ms.value)

    def eval(self, xyz) -> np.ndarray:
        """(n, 4): distance and its gradient at the points."""
        p = np.ascontiguousarray(xyz, np.float32).reshape(-1, 3)
        out = np.empty((p.shape[0], 4), np.float32)
        _check(lib.sph_mlp_eval(self._h, as_fptr(p), p.shape[0], as_fptr(out)), self._h)
        return out


def default_neural_config(desc: SceneDesc, tolerance_h: float = 0.05, band_h: float = 2.0, brick: int = 4,
                          normal_tolerance: float = 0.05):
    """Validation bars per brick: RMS distance error tolerance_h * h over
    samples whose label exceeds -band_h * h, and RMS wall-normal error
    normal_tolerance (|grad d_network - grad d_grid|, world units) over
    samples in the wall-penalty zone."""
    h = float(desc.h)
    cfg = SphNeuralConfig()
    cfg.tolerance = tolerance_h * h
    cfg.band = band_h * h
    cfg.normal_tolerance = float(normal_tolerance)
    cfg.brick = int(brick)
    return cfg


def validation_samples(desc: SceneDesc, n: int = 1 << 20, seed: int = 1234, maps=None) -> np.ndarray:
    """Held-out grid-labelled samples: 80% within 3 h of the surface, 20%
    uniform."""
    return grid_samples(desc, n, near_frac=0.8, near_width=3.0 * float(desc.h), seed=seed, maps=maps)


def report_dict(r: SphNeuralReport) -> dict:
    return {f: (float(getattr(r, f)) if "error" in f else int(getattr(r, f))) for f, _ in SphNeuralReport._fields_}


def attach(target, net: NeuralSDF, desc: SceneDesc, scene: int = 0, val=None, cfg=None, maps=None) -> dict:
    """Attaches the network as the distance source of a ParticleSimulation
    (or of scene `scene` of a Batch); returns the attach report.  `maps`:
    the scene's map obstacles, for the default validation samples."""
    from .simulation import Batch
    if val is None:
        val = validation_samples(desc, maps=maps)
    val = np.ascontiguousarray(val, np.float32).reshape(-1, 4)
    cfg = cfg or default_neural_config(desc)
    rep = SphNeuralReport()
    if isinstance(target, Batch):
        st = lib.sph_batch_attach_neural_boundary(target._b, int(scene), net.handle, as_fptr(val), val.shape[0],
                                                  C.byref(cfg), C.byref(rep))
        if st != Status.OK:
            msg = lib.sph_batch_last_error_message(target._b)
            raise SphError(st, (msg or b"").decode(errors="replace"))
    else:
        h = target._need()
        st = lib.sph_attach_neural_boundary(h, net.handle, as_fptr(val), val.shape[0], C.byref(cfg), C.byref(rep))
        if st != Status.OK:
            msg = lib.sph_last_error_message(h)
            raise SphError(st, (msg or b"").decode(errors="replace"))
    target._neural = net          # keep the network alive while attached
    return report_dict(rep)


def train_for_scene(desc: SceneDesc, steps: int = 20000, batch: int = 65536, n_samples: int = 1 << 22,
                    seed: int = 1, maps=None) -> NeuralSDF:
    """The configuration's training recipe: 4M samples of the boundary grid's
    distance texture, 70% uniform / 30% within 3 h of the surface; output
    scaled by r0 (the texture's range); Adam lr 1e-3 with cosine decay to
    1e-5.  `maps`: the scene's map obstacles (scenes.mesh_obstacle)."""
    h = float(desc.h)
    net = NeuralSDF.for_scene(desc, seed=seed, out_scale=float(desc.rest_distance), lr=1e-3, lr_min=1e-5,
                              decay_steps=steps)
    net.set_samples(grid_samples(desc, n_samples, near_frac=0.3, near_width=3.0 * h, seed=seed + 1, maps=maps))
    net.train(steps, batch)
    return net


def boundary_map_samples(bm: dict) -> np.ndarray:
    """(n, 4) float32 (x, y, z, distance) at the nodes of a generated boundary
    map (cudasphnew_amd.meshsdf.generate) that hold a distance (present), the
    band within max(restDistance, compactSupport) of the mesh."""
    dims = bm["dims"]
    lo = np.asarray(bm["lo"], np.float32)
    hi = np.asarray(bm["hi"], np.float32)
    iz, iy, ix = np.nonzero(bm["present"])
    sp = (hi - lo) / (np.asarray(dims, np.float32) - 1)
    out = np.empty((ix.shape[0], 4), np.float32)
    out[:, 0] = lo[0] + ix * sp[0]
    out[:, 1] = lo[1] + iy * sp[1]
    out[:, 2] = lo[2] + iz * sp[2]
    out[:, 3] = bm["distance"][iz, iy, ix]
    return out


def train_for_boundary_map(bm: dict, band: float, steps: int = 20000, batch: int = 65536, seed: int = 1,
                           samples: np.ndarray | None = None) -> NeuralSDF:
    """The learned SDF of a mesh: trained on the band nodes of its generated
    boundary map (or on given samples of them), over the map's box, with the
    band-compressed target tanh(d / band); Adam lr 1e-3 with cosine decay to
    1e-5 (the recipe of train_for_scene)."""
    net = NeuralSDF(bm["lo"], bm["hi"], seed=seed, out_scale=band, lr=1e-3, lr_min=1e-5, decay_steps=steps,
                    band=band)
    net.set_samples(boundary_map_samples(bm) if samples is None else samples)
    net.train(steps, batch)
    return net
