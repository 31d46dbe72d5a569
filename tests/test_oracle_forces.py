"""Pins the oracle's acceleration pair terms, colour field and surface
tension (coarse and sub-particle) without using its formulas: the expected
accelerations are assembled from the SPH kernel *functions* of Mueller,
Charypar and Gross 2003 ("Particle-Based Fluid Simulation for Interactive
Applications", eqs. 20-22: W_poly6, W_spiky, W_viscosity), whose gradients
and Laplacians are taken here by finite differences, and from the force
equations they define:

  rho_i = sum_j m W_poly6(x_i - x_j)                  (self included)
  p_i   = k (rho_i - rho0)
  a_p   = -sum_{j!=i} m (p_i/rho_i^2 + p_j/rho_j^2) grad W_spiky(x_i - x_j)
          (Monaghan 1992's symmetric pressure gradient, the form of
           particle_simulation.cu:1516-1525)
  a_v   = (mu/rho_i) sum_{j!=i} m (v_j - v_i)/rho_j lapl W_visc(x_i - x_j)
  c(x)  = sum_{j!=i} m/rho_j W_poly6(x - x_j)         (colour field)
  a_s   = -(sigma/rho_i) lapl c(x_i) grad c(x_i)/|grad c(x_i)|
          when |grad c(x_i)| > normThresh, no wall term active
          (particle_simulation.cu:455-466; sub: 551-560)
  a     = a_p + a_v + a_s - g y

The reference's kernel coefficients (cu:1779-1788) are the closed-form
derivatives of these functions, rounded to FP32 (so results agree to ~1e-7
relative, and the bars are 1e-6); the finite differences make the test
independent of them, so a spiky gradient with (h - r)^3, a viscosity
Laplacian with (h - r)^2, a dropped colour-Laplacian factor, a tension with
the wrong sign or factor, or sub coefficients with the wrong power of two
each fail a check.  Sub-particles use the same functions with h_sub.
"""
import ctypes
import math

import numpy as np
import pytest

import cudasphnew_amd as S
from cudasphnew_amd import Field
from oracle import oracle as O


# ---- the kernel functions (Mueller et al. 2003, eqs. 20-22) -----------------------

def w_poly6(r, h):
    """W_poly6 inside its support (r <= h; the tests select the pairs)."""
    r = np.asarray(r, np.float64)
    return 315.0 / (64.0 * math.pi * h ** 9) * (h * h - r * r) ** 3


def w_spiky(r, h):
    r = np.asarray(r, np.float64)
    return 15.0 / (math.pi * h ** 6) * (h - r) ** 3


def w_visc(r, h):
    r = np.asarray(r, np.float64)
    return 15.0 / (2.0 * math.pi * h ** 3) * (-r ** 3 / (2 * h ** 3) + r * r / (h * h) + h / (2 * r) - 1.0)


def fd_grad(f, x, eps):
    """Central difference gradient of a scalar field f(points[k,3]) at x."""
    g = np.zeros(3)
    for a in range(3):
        e = np.zeros(3)
        e[a] = eps
        g[a] = (f((x + e)[None])[0] - f((x - e)[None])[0]) / (2 * eps)
    return g


def fd_lapl(f, x, eps):
    """Laplacian by central second differences, Richardson-extrapolated
    (eps and eps/2) to fourth order."""
    def lap(e_):
        s = 0.0
        f0 = f(x[None])[0]
        for a in range(3):
            e = np.zeros(3)
            e[a] = e_
            s += (f((x + e)[None])[0] - 2 * f0 + f((x - e)[None])[0]) / (e_ * e_)
        return s
    return (4 * lap(eps / 2) - lap(eps)) / 3


def field_of(kernel, centers, weights, h):
    """x -> sum_j weights_j kernel(|x - c_j|) over a fixed neighbour set."""
    def f(pts):
        d = np.linalg.norm(pts[:, None, :] - centers[None, :, :], axis=2)
        return (kernel(d, h) * weights[None, :]).sum(1)
    return f


def expected_accel(pos, vel, m, h, k, rho0, mu, sigma, thresh, g):
    """Density, pressure and acceleration of every particle from the
    definitions above (no wall term)."""
    n = pos.shape[0]
    d = np.linalg.norm(pos[:, None] - pos[None], axis=2)
    # the kernels are evaluated inside their support only (pairs with
    # r <= h), so the finite differences never straddle the cutoff
    rho = m * np.where(d <= h, w_poly6(np.minimum(d, h), h), 0.0).sum(1)
    p = k * (rho - rho0)
    acc = np.zeros((n, 3))
    cgn = np.zeros(n)
    eps = 1e-4 * h
    for i in range(n):
        others = [j for j in range(n) if j != i and d[i, j] <= h]
        a = np.zeros(3)
        for j in others:
            gw = fd_grad(lambda q: w_spiky(np.linalg.norm(q, axis=1), h), pos[i] - pos[j], eps)
            a -= m * (p[i] / rho[i] ** 2 + p[j] / rho[j] ** 2) * gw
            lw = fd_lapl(lambda q: w_visc(np.linalg.norm(q, axis=1), h), pos[i] - pos[j], 20 * eps)
            a += mu / rho[i] * m * (vel[j] - vel[i]) / rho[j] * lw
        if others:
            oj = np.array(others)
            c = field_of(w_poly6, pos[oj], m / rho[oj], h)
            gc = fd_grad(c, pos[i], eps)
            lc = fd_lapl(c, pos[i], 20 * eps)
            cgn[i] = np.linalg.norm(gc)
            if cgn[i] > thresh:
                a -= sigma * lc * gc / cgn[i] / rho[i]
        a[1] -= g
        acc[i] = a
    return rho, p, acc, cgn


@pytest.fixture(scope="module")
def base():
    return S.make_scene("dam16k")


def free_space(sc, params=None):
    """The scene's grid geometry with a boundary far from every wall
    (distance -r0 everywhere: no wall term, no collision)."""
    shape = sc.distance.shape
    r0 = sc.desc.rest_distance
    prm = S.SphParams.from_buffer_copy(sc.params if params is None else params)
    prm.collision = 0
    z = np.zeros(shape, np.float32)
    return S.Scene(sc.desc, prm, sc.positions, np.full(shape, -r0, np.float32), z, z)


def _consts(p, sub=False):
    h = float(p.compact_support_sub if sub else p.compact_support)
    m = float(p.sub_particle_mass if sub else p.particle_mass)
    return dict(m=m, h=h, k=float(p.gas_stiffness), rho0=float(p.rest_density), mu=float(p.dynamic_viscosity),
                sigma=float(p.tension_coefficient), thresh=float(p.norm_thresh), g=float(p.gravity))


def _check(acc, exp, rel):
    scale = max(np.abs(exp).max(), 1e-12)
    err = np.abs(acc - exp).max() / scale
    assert err <= rel, f"acceleration differs from the kernel-function form by {err:.3e} (relative)"


# ---- coarse particles --------------------------------------------------------------

def test_pair_terms_from_kernel_functions(base):
    """Three particles within h of each other (not collinear), distinct
    velocities, tension off: pressure + viscosity against the kernel
    functions' finite-difference derivatives."""
    sc = free_space(base)
    sc.params.norm_thresh = 1e30
    sc.params.gravity = 0.0
    h = float(sc.params.compact_support)
    pos = np.array([[0.30, 0.30, 0.30],
                    [0.30 + 0.45 * h, 0.30 + 0.10 * h, 0.30],
                    [0.30 + 0.05 * h, 0.30 + 0.55 * h, 0.30 + 0.30 * h]])
    vel = np.array([[0.2, -0.1, 0.05], [-0.3, 0.4, 0.0], [0.1, 0.0, -0.25]])
    orc = O.OracleSim(sc, real=0, positions=pos, velocities=vel)
    orc.step(1)
    c = _consts(sc.params)
    rho, p, exp, _ = expected_accel(pos, vel, **c)
    assert np.allclose(orc.read(Field.DENSITY), rho, rtol=1e-6)
    assert np.allclose(orc.read(Field.PRESSURE), p, rtol=1e-6)
    _check(orc.read(Field.ACCEL), exp, 1e-6)


def test_pressure_alone_and_viscosity_alone(base):
    """Each pair term separately (the other switched off by its parameter),
    so a compensating pair of mistakes cannot pass."""
    h = float(base.params.compact_support)
    pos = np.array([[0.30, 0.30, 0.30], [0.30 + 0.6 * h, 0.30 + 0.2 * h, 0.30 - 0.1 * h],
                    [0.30 - 0.3 * h, 0.30 + 0.4 * h, 0.30 + 0.2 * h]])
    vel = np.array([[0.5, 0.0, 0.0], [-0.5, 0.2, 0.1], [0.0, -0.3, 0.4]])
    for stiff, mu in ((3.0, 0.0), (0.0, 3.5)):
        sc = free_space(base)
        sc.params.norm_thresh = 1e30
        sc.params.gravity = 0.0
        sc.params.gas_stiffness = stiff
        sc.params.rest_density = 1.0          # positive pressure: repulsive
        sc.params.dynamic_viscosity = mu
        orc = O.OracleSim(sc, real=0, positions=pos, velocities=vel)
        orc.step(1)
        _, _, exp, _ = expected_accel(pos, vel, **_consts(sc.params))
        assert np.abs(exp).max() > 1.0
        _check(orc.read(Field.ACCEL), exp, 1e-6)


def test_surface_tension_from_colour_field(base):
    """A small cluster, so that |grad c| exceeds normThresh for every
    particle: the tension term -sigma lapl(c) grad(c)/|grad(c)| with the
    colour field's gradient and Laplacian taken by finite differences of
    c(x) = sum_j m/rho_j W_poly6(x - x_j)."""
    sc = free_space(base)
    sc.params.gas_stiffness = 0.0         # isolate tension (+ viscosity, zero here)
    h = float(sc.params.compact_support)
    rng = np.random.default_rng(11)
    pos = 0.3 + h * np.array([[0, 0, 0], [0.5, 0.1, 0.0], [0.1, 0.45, 0.2], [-0.3, 0.2, 0.35],
                              [0.25, -0.35, 0.3]]) + 1e-4 * h * rng.standard_normal((5, 3))
    vel = np.zeros_like(pos)
    orc = O.OracleSim(sc, real=0, positions=pos, velocities=vel)
    orc.step(1)
    c = _consts(sc.params)
    rho, _, exp, cgn = expected_accel(pos, vel, **c)
    # every particle takes the tension branch, away from the switch
    assert np.all(cgn > 1.2 * c["thresh"]), cgn
    ocgn = orc.branch_margin()   # the oracle agrees it is on that branch
    assert np.all(ocgn > 0.1)
    # the tension term is a real part of the result
    _, _, exp_no_tension, _ = expected_accel(pos, vel, **dict(c, sigma=0.0))
    assert np.abs(exp - exp_no_tension).max() > 0.05 * np.abs(exp).max()
    _check(orc.read(Field.ACCEL), exp, 1e-6)
    # the colour-gradient norm the oracle compared with normThresh
    got = np.empty(pos.shape[0])
    assert O.lib.orc_read(orc._s, O.FIELD_COLOR_GRAD_NORM, got.ctypes.data_as(ctypes.c_void_p), got.nbytes) == 5
    assert np.allclose(got, cgn, rtol=1e-6)


def test_all_terms_together_with_tension_and_gravity(base):
    sc = free_space(base)
    h = float(sc.params.compact_support)
    pos = 0.3 + h * np.array([[0, 0, 0], [0.55, 0.05, -0.1], [0.1, 0.5, 0.2], [-0.35, 0.25, 0.3]])
    vel = np.array([[0.1, 0.2, 0.0], [-0.2, 0.0, 0.1], [0.0, -0.1, 0.3], [0.3, 0.1, -0.2]])
    orc = O.OracleSim(sc, real=0, positions=pos, velocities=vel)
    orc.step(1)
    rho, _, exp, cgn = expected_accel(pos, vel, **_consts(sc.params))
    assert np.all(np.abs(cgn / float(sc.params.norm_thresh) - 1) > 0.1)
    _check(orc.read(Field.ACCEL), exp, 1e-6)


def test_finite_differences_see_the_plausible_mistakes(base):
    """The finite-difference forms are not the oracle's closed forms: check
    that the mistakes this file guards against change the expected values by
    far more than the tolerance (so the tests above would fail them)."""
    h = 0.02
    r = np.array([0.3 * h, 0.0, 0.4 * h])
    rn = np.linalg.norm(r)
    gs = fd_grad(lambda q: w_spiky(np.linalg.norm(q, axis=1), h), r, 1e-4 * h)
    right = -45.0 / (math.pi * h ** 6) * (h - rn) ** 2 * r / rn
    wrong = -45.0 / (math.pi * h ** 6) * (h - rn) ** 3 * r / rn
    assert np.allclose(gs, right, rtol=1e-7)
    assert np.abs(gs - wrong).max() > 0.1 * np.abs(right).max()
    lv = fd_lapl(lambda q: w_visc(np.linalg.norm(q, axis=1), h), r, 2e-3 * h)
    assert lv == pytest.approx(45.0 / (math.pi * h ** 6) * (h - rn), rel=1e-7)
    lp = fd_lapl(lambda q: w_poly6(np.linalg.norm(q, axis=1), h), r, 2e-3 * h)
    full = -945.0 / (32 * math.pi * h ** 9) * (h * h - rn * rn) * (3 * h * h - 7 * rn * rn)
    dropped = -945.0 / (32 * math.pi * h ** 9) * (h * h - rn * rn)
    assert lp == pytest.approx(full, rel=1e-7)
    assert abs(lp - dropped) > 0.1 * abs(full)


# ---- sub-particles (compute_sub_particle_acceleration, cu:477-568, 1558-1643) ------

def _two_scale_subs(base, pos, vel, **over):
    prm = S.SphParams.from_buffer_copy(base.params)
    prm.sub_placement = 1                   # children at +-h/4 around the parent
    for a in range(3):                      # every particle splits
        prm.split_lo[a] = -1e30
        prm.split_hi[a] = 1e30
    for k, v in over.items():
        setattr(prm, k, v)
    sc = free_space(base, prm)
    orc = O.OracleSim(sc, real=0, positions=pos, velocities=vel)
    orc.step(1, two_scale=True)
    cn = orc.counts()
    assert cn.num_split == pos.shape[0] and cn.num_boundary == 0
    xs = orc.read(Field.SUB_POSITION)
    vs = orc.read(Field.SUB_VELOCITY)
    as_ = orc.read(Field.SUB_ACCEL)
    dts = float(np.float32(prm.time_step_sub))
    # the sub-particles were integrated after their acceleration pass
    # (v += dt a; x += dt v): undo it to get the state the pass saw
    x0 = xs - dts * vs
    v0 = vs - dts * as_
    return sc, orc, x0, v0, as_


def test_sub_particle_forces_from_kernel_functions(base):
    """Two split parents 0.6 h apart with different velocities give 16
    children at +-h/4; children of different parents interact within
    h_sub = h/2.  Sub density, pressure and acceleration (pressure,
    viscosity, colour field, tension) against the kernel functions at
    h_sub with the sub-particle mass."""
    h = float(base.params.compact_support)
    pos = np.array([[0.30, 0.30, 0.30], [0.30 + 0.6 * h, 0.30 + 0.1 * h, 0.30 - 0.05 * h]])
    vel = np.array([[0.3, -0.2, 0.1], [-0.25, 0.15, 0.0]])
    sc, orc, x0, v0, acc = _two_scale_subs(base, pos, vel)
    p = sc.params
    c = _consts(p, sub=True)
    assert c["h"] == pytest.approx(h / 2, rel=1e-7)
    rho, pr, exp, cgn = expected_accel(x0, v0, **c)
    assert np.allclose(orc.read(Field.SUB_DENSITY), rho, rtol=1e-6)
    assert np.allclose(orc.read(Field.SUB_PRESSURE), pr, rtol=1e-6)
    near = np.abs(cgn / c["thresh"] - 1) > 1e-3
    assert near.all()
    assert (cgn > c["thresh"]).any()          # the tension branch is exercised
    _check(acc, exp, 1e-6)


def test_sub_particle_pair_terms_without_tension(base):
    h = float(base.params.compact_support)
    pos = np.array([[0.30, 0.30, 0.30], [0.30 + 0.55 * h, 0.30 - 0.1 * h, 0.30 + 0.1 * h],
                    [0.30 + 0.2 * h, 0.30 + 0.5 * h, 0.30]])
    vel = np.array([[0.3, 0.0, 0.1], [-0.2, 0.25, 0.0], [0.0, -0.3, 0.2]])
    for stiff, mu in ((3.0, 0.0), (0.0, 3.5)):
        sc, orc, x0, v0, acc = _two_scale_subs(base, pos, vel, norm_thresh=1e30, gravity=0.0,
                                               gas_stiffness=stiff, rest_density=1.0, dynamic_viscosity=mu)
        _, _, exp, _ = expected_accel(x0, v0, **_consts(sc.params, sub=True))
        assert np.abs(exp).max() > 1.0
        _check(acc, exp, 1e-6)
