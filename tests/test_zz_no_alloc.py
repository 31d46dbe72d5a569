"""sph_step_noalloc -- the allocation-free plain step over caller device
memory -- through the C ABI, against sph_step on the same scene: positions,
velocities and densities bit for bit (the same kernels, the same default
variants), on two calls in a row, the second with fewer particles (the
shared scratch array keeps the first call's contents).  The caller's device
buffers and stream come from the HIP runtime the library itself links
(ctypes on its libamdhip64), as a caller of the C ABI would make them."""
import ctypes as C

import numpy as np
import pytest

import cudasphnew_amd as S
from cudasphnew_amd import Field
from cudasphnew_amd import _native as N

pytestmark = pytest.mark.gpu


def _hip():
    """The HIP runtime libsph_hip.so is linked against (not another copy a
    framework may have loaded into the process)."""
    path = None
    for line in open("/proc/self/maps"):
        if "libamdhip64" in line and line.split()[-1].startswith("/opt/rocm"):
            path = line.split()[-1]
            break
    hip = C.CDLL(path or "/opt/rocm/lib/libamdhip64.so")
    hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    hip.hipFree.argtypes = [C.c_void_p]
    hip.hipStreamCreate.argtypes = [C.POINTER(C.c_void_p)]
    hip.hipStreamSynchronize.argtypes = [C.c_void_p]
    hip.hipStreamDestroy.argtypes = [C.c_void_p]
    return hip


class DevBuf:
    """Device copy of a host array (hipMalloc / hipMemcpy / hipFree)."""

    def __init__(self, hip, host: np.ndarray):
        self.hip, self.nbytes = hip, host.nbytes
        self.ptr = C.c_void_p()
        assert hip.hipMalloc(C.byref(self.ptr), self.nbytes) == 0
        assert hip.hipMemcpy(self.ptr, host.ctypes.data, self.nbytes, 1) == 0     # host to device

    def get(self, like: np.ndarray) -> np.ndarray:
        out = np.empty_like(like)
        assert self.hip.hipMemcpy(out.ctypes.data, self.ptr, self.nbytes, 2) == 0  # device to host
        return out

    def free(self):
        self.hip.hipFree(self.ptr)


def _device_boundary(sc, hip):
    sb = sc.sph_boundary()
    grid = np.ascontiguousarray(
        np.stack([sc.distance, sc.density, sc.viscosity, np.zeros_like(sc.distance)], -1).reshape(-1, 4),
        dtype=np.float32)
    g = DevBuf(hip, grid)
    d = DevBuf(hip, np.ascontiguousarray(sc.distance.reshape(-1), dtype=np.float32))
    b = N.SphDeviceBoundary()
    for a in range(3):
        b.origin[a] = sb.origin[a]
        b.dims[a] = sb.dims[a]
    b.spacing = sb.spacing
    b.rest_distance = sb.rest_distance
    b.grid = g.ptr.value
    b.distance = d.ptr.value
    return b, (g, d)


def _noalloc(sc, steps, hip, stream, flags=None):
    b, keep = _device_boundary(sc, hip)
    n = sc.n
    p4 = np.zeros((n, 4), np.float32)
    p4[:, :3] = sc.positions
    v4 = np.zeros((n, 4), np.float32)
    pos, vel = DevBuf(hip, p4), DevBuf(hip, v4)
    st = N.lib.sph_step_noalloc(C.byref(sc.params), C.byref(b), n, pos.ptr, vel.ptr, steps,
                                None if flags is None else flags.ptr, stream)
    assert st == 0, N.lib.sph_last_error_message(None)
    assert hip.hipStreamSynchronize(stream) == 0
    out = pos.get(p4), vel.get(v4)
    for x in (pos, vel, *keep):
        x.free()
    return out


def _handle(sc, steps):
    sim = S.ParticleSimulation(sc).Init()
    sim.Advance(steps)
    sim.sync()
    out = sim.read(Field.POSITION), sim.read(Field.VELOCITY), sim.read(Field.DENSITY)
    sim.close()
    return out


def test_noalloc_step_equals_handle_step_on_two_calls():
    hip = _hip()
    stream = C.c_void_p()
    assert hip.hipStreamCreate(C.byref(stream)) == 0
    base = S.make_scene("dam16k")
    flags = DevBuf(hip, np.ones(1, np.uint32))
    # first call: the whole 16K dam break, 3 steps
    pos, vel = _noalloc(base, 3, hip, stream, flags)
    x, v, rho = _handle(base, 3)
    assert np.array_equal(pos[:, :3], x) and np.array_equal(vel[:, :3], v)
    assert np.array_equal(pos[:, 3], rho) and np.array_equal(vel[:, 3], rho)
    assert int(flags.get(np.zeros(1, np.uint32))[0]) == 0
    # second call, smaller: 10,007 particles of the same block, 2 steps
    small = base.with_positions(base.positions[:10007].copy())
    pos2, vel2 = _noalloc(small, 2, hip, stream, flags)
    x2, v2, rho2 = _handle(small, 2)
    assert np.array_equal(pos2[:, :3], x2) and np.array_equal(vel2[:, :3], v2)
    assert np.array_equal(pos2[:, 3], rho2)
    assert int(flags.get(np.zeros(1, np.uint32))[0]) == 0
    # the particles moved
    assert np.abs(pos2[:, :3] - small.positions).max() > 0
    flags.free()
    hip.hipStreamDestroy(stream)


def test_noalloc_step_rejects_sizes_beyond_its_limits():
    hip = _hip()
    sc = S.make_scene("dam16k")
    b, keep = _device_boundary(sc, hip)
    buf = DevBuf(hip, np.zeros((8, 4), np.float32))
    # more particles than the scratch array holds (checked before any launch)
    st = N.lib.sph_step_noalloc(C.byref(sc.params), C.byref(b), N_MAX + 1, buf.ptr, buf.ptr, 1, None, None)
    assert st == S.Status.INVALID_ARG
    assert b"SPH_NOALLOC_MAX_PARTICLES" in N.lib.sph_last_error_message(None)
    # a cell grid larger than the scratch array's cell table (the 1M scene's)
    big = S.make_scene("dam1m")
    st = N.lib.sph_step_noalloc(C.byref(big.params), C.byref(b), 8, buf.ptr, buf.ptr, 1, None, None)
    assert st == S.Status.INVALID_ARG
    assert b"SPH_NOALLOC_MAX_CELLS" in N.lib.sph_last_error_message(None)
    for x in (buf, *keep):
        x.free()


N_MAX = 262144
