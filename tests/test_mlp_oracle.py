"""Pins the FP64 MLP restatement (oracle/mlp.py) on the CPU: closed forms of
the features, finite-difference checks of its back-propagation, bf16
rounding known answers and Adam's first-step property."""
import numpy as np
import pytest

from oracle import mlp as M


def _rng_u(n, seed=0):
    return np.random.default_rng(seed).uniform(-1, 1, (n, 3))


def test_features_closed_form():
    u = _rng_u(50)
    g = M.features(u)
    assert g.shape == (50, M.K1)
    assert np.array_equal(g[:, :3], u)
    assert np.all(g[:, M.NFEAT:] == 0.0)
    for k in range(M.NF):
        for a in range(3):
            s = g[:, 3 + 6 * k + 2 * a]
            c = g[:, 3 + 6 * k + 2 * a + 1]
            assert np.allclose(s * s + c * c, 1.0, atol=1e-14)
            assert np.allclose(s, np.sin(np.pi * M.FREQS[k] * u[:, a]))
    g0 = M.features(np.zeros((1, 3)))
    assert np.allclose(g0[0, 3:M.NFEAT:2], 0.0) and np.allclose(g0[0, 4:M.NFEAT:2], 1.0)
    assert M.FREQS[0] == 1.0 and M.FREQS[15] == np.float32(128.0)


def test_normalized_maps_box_to_unit_cube():
    lo, hi = np.array([-0.1, 0.0, 0.2], np.float32), np.array([0.5, 1.0, 0.4], np.float32)
    u = M.normalized(np.stack([lo, hi, 0.5 * (lo + hi)]), lo, hi)
    assert np.allclose(u, [[-1, -1, -1], [1, 1, 1], [0, 0, 0]], atol=1e-6)


def test_param_layout():
    assert M.NUM_PARAMS == 16641
    th = np.arange(M.NUM_PARAMS, dtype=np.float64)
    assert np.array_equal(M.pack(M.unpack(th)), th)


@pytest.mark.parametrize("seed", [1, 2])
def test_weight_gradient_matches_finite_differences(seed):
    th = M.xavier_init(seed)
    u = _rng_u(24, seed)
    yt = np.random.default_rng(seed + 9).normal(0, 0.3, 24)
    loss, g = M.loss_and_grad(th, u, yt)
    rng = np.random.default_rng(seed)
    offs = [M.O_W1, M.O_B1, M.O_W2, M.O_B2, M.O_W3, M.O_B3, M.O_W4, M.O_B4, M.NUM_PARAMS]
    idx = []
    for a, b in zip(offs[:-1], offs[1:]):
        hi = b if a != M.O_W1 else M.O_W1 + M.NFEAT * M.H
        idx += list(rng.integers(a, hi, 6))
    eps = 1e-6
    for i in idx:
        tp, tm = th.copy(), th.copy()
        tp[i] += eps
        tm[i] -= eps
        fd = (M.loss_and_grad(tp, u, yt)[0] - M.loss_and_grad(tm, u, yt)[0]) / (2 * eps)
        assert abs(fd - g[i]) <= 1e-7 + 1e-5 * abs(fd), (i, fd, g[i])
    # padding rows of W1 see zero features: zero gradient
    assert np.all(g[M.O_W1 + M.NFEAT * M.H:M.O_B1] == 0.0)


def test_input_gradient_matches_finite_differences():
    th = M.xavier_init(4)
    u = _rng_u(10, 4) * 0.9
    y, gu = M.input_grad(th, u)
    eps = 1e-7
    for a in range(3):
        up, um = u.copy(), u.copy()
        up[:, a] += eps
        um[:, a] -= eps
        fd = (M.forward(th, up)[0] - M.forward(th, um)[0]) / (2 * eps)
        assert np.allclose(fd, gu[:, a], rtol=1e-5, atol=1e-6)


def test_bf16_rounding_known_answers():
    x = np.array([1.0, 1 + 2.0 ** -8, 1 + 3 * 2.0 ** -8, -2.5, 1 + 2.0 ** -8 + 2.0 ** -20, 0.0])
    r = M.bf16(x)
    assert np.array_equal(r, [1.0, 1.0, 1 + 2.0 ** -6, -2.5, 1 + 2.0 ** -7, 0.0])


def test_adam_first_step_moves_by_lr():
    th = np.zeros(5)
    g = np.array([1e-3, -2.0, 5.0, -1e-5, 3e-2])
    th1, m, v = M.adam(th, np.zeros(5), np.zeros(5), g, 1, lr=1e-3)
    # first step: m_hat = g, v_hat = g^2 -> step = lr * g / (|g| + eps)
    assert np.allclose(th1, -1e-3 * g / (np.abs(g) + 1e-8), rtol=1e-12)
    assert np.allclose(m, 0.1 * g) and np.allclose(v, 0.001 * g * g)


def test_bf16_model_changes_little():
    th = M.xavier_init(3)
    u = _rng_u(200, 3)
    y, _ = M.forward(th, u)
    yb, _ = M.forward(th, u, bf=True)
    assert np.max(np.abs(y - yb)) < 0.05 * np.max(np.abs(y)) + 1e-3
