"""Learned boundary SDF on the GPU against the FP64 oracle (oracle/mlp.py).

Bars: forward values and input gradients against the oracle that rounds the
product operands to bfloat16 where the kernels do (FP32 accumulation vs FP64
sums and rare one-ulp bfloat16 flips of intermediate activations leave
~1e-6 typical, ~1e-3 worst-case differences); the weight gradient of one
batch per tensor within 2e-3 (relative Frobenius norm) of that oracle and
within 5e-2 of the pure FP64 gradient; one Adam step to 1e-7 absolute.
"""
import numpy as np
import pytest

import cudasphnew_amd as S
from oracle import mlp as M
from oracle import oracle as O

pytestmark = pytest.mark.gpu

LO = np.array([-0.05, -0.02, -0.04], np.float32)
HI = np.array([0.69, 0.66, 0.7], np.float32)


@pytest.fixture(scope="module")
def net():
    n = S.NeuralSDF(LO, HI, seed=5)
    th = M.xavier_init(11).astype(np.float32)
    n.set_params(th)
    return n, th.astype(np.float64)


def _points(n, seed):
    rng = np.random.default_rng(seed)
    return (LO + rng.random((n, 3)) * (HI - LO)).astype(np.float32)


def test_params_roundtrip_and_init():
    n = S.NeuralSDF(LO, HI, seed=3)
    th = n.get_params()
    assert th.shape == (M.NUM_PARAMS,)
    p = M.unpack(th)
    assert np.all(p["W1"][M.NFEAT:] == 0.0) and np.all(p["b1"] == 0.0)
    lim = np.sqrt(6.0 / (M.NFEAT + M.H))
    # U(-lim, lim): std lim / sqrt(3)
    assert np.abs(p["W1"]).max() <= lim and p["W1"][:M.NFEAT].std() > 0.5 * lim
    n2 = S.NeuralSDF(LO, HI, seed=3)
    assert np.array_equal(n2.get_params(), th)


def test_forward_and_input_gradient_match_oracle(net):
    n, th = net
    x = _points(3000, 1)
    out = n.eval(x)
    u = M.normalized(x, LO, HI)
    scale = float(n.config.out_scale)
    y_b, gu_b = M.input_grad(th, u, bf=True)
    d_b, _ = M.decode(y_b, 0.0, scale)
    err = np.abs(out[:, 0] - d_b)
    print(f"forward: median {np.median(err):.3g}, max {err.max():.3g} (|d| ~ {np.abs(d_b).mean():.3g})")
    assert np.median(err) <= 1e-5 * (1 + np.abs(d_b).mean())
    assert err.max() <= 2e-3 * (1 + np.abs(d_b).max())
    gx_b = scale * gu_b * (2.0 / (HI - LO).astype(np.float64))
    gerr = np.linalg.norm(out[:, 1:] - gx_b, axis=1) / (np.linalg.norm(gx_b, axis=1) + 1e-3)
    print(f"input gradient: median {np.median(gerr):.3g}, max {gerr.max():.3g}")
    assert np.median(gerr) <= 1e-4 and gerr.max() <= 2e-2
    # the pure FP64 network differs by the bfloat16 rounding only
    y64, _ = M.forward(th, u)
    assert np.max(np.abs(out[:, 0] - scale * y64)) <= 0.05 * (1 + np.abs(scale * y64).max())


def test_weight_gradient_matches_oracle(net):
    n, th = net
    B = 4096
    x = _points(B, 2)
    d = (0.3 * np.sin(3 * x[:, 0]) - 0.2 * x[:, 1] + 0.1).astype(np.float32)
    n.set_samples(np.column_stack([x, d]))
    loss, g = n.gradient(0, B)
    u = M.normalized(x, LO, HI)
    yt = M.target(d, 0.0, float(n.config.out_scale))
    loss_b, g_b = M.loss_and_grad(th, u, yt, bf=True)
    loss_64, g_64 = M.loss_and_grad(th, u, yt)
    assert loss == pytest.approx(loss_b, rel=1e-4)
    offs = [("W1", M.O_W1, M.O_B1), ("b1", M.O_B1, M.O_W2), ("W2", M.O_W2, M.O_B2), ("b2", M.O_B2, M.O_W3),
            ("W3", M.O_W3, M.O_B3), ("b3", M.O_B3, M.O_W4), ("w4", M.O_W4, M.O_B4), ("b4", M.O_B4, M.NUM_PARAMS)]
    for name, a, b in offs:
        e_b = np.linalg.norm(g[a:b] - g_b[a:b]) / np.linalg.norm(g_b[a:b])
        e_64 = np.linalg.norm(g[a:b] - g_64[a:b]) / np.linalg.norm(g_64[a:b])
        print(f"{name}: vs bf16 oracle {e_b:.3g}, vs FP64 {e_64:.3g}")
        assert e_b <= 2e-3, name
        assert e_64 <= 5e-2, name


def test_one_adam_step_matches_oracle(net):
    n, th = net
    B = 2048
    x = _points(B, 3)
    d = (0.1 - 0.4 * x[:, 2]).astype(np.float32)
    n.set_params(th.astype(np.float32))
    n.set_samples(np.column_stack([x, d]))
    _, g = n.gradient(0, B)
    n.train(1, B)
    th1 = n.get_params().astype(np.float64)
    cfg = n.config
    ref, _, _ = M.adam(th.astype(np.float32).astype(np.float64), np.zeros_like(th), np.zeros_like(th),
                       g.astype(np.float64), 1, lr=cfg.lr, b1=cfg.beta1, b2=cfg.beta2, eps=cfg.eps)
    assert np.max(np.abs(th1 - ref)) <= 1e-7
    n.set_params(th.astype(np.float32))


def test_band_target_and_decoding_match_oracle():
    band = 0.05
    n = S.NeuralSDF(LO, HI, seed=5, band=band)
    th = M.xavier_init(12)
    th[M.O_W4:M.O_B4] *= 0.3          # keep |y| < 1
    n.set_params(th.astype(np.float32))
    th = th.astype(np.float32).astype(np.float64)
    x = _points(2048, 4)
    d = (0.12 - 0.3 * x[:, 1]).astype(np.float32)
    n.set_samples(np.column_stack([x, d]))
    loss, g = n.gradient(0, 2048)
    u = M.normalized(x, LO, HI)
    loss_b, g_b = M.loss_and_grad(th, u, M.target(d, band, 1.0), bf=True)
    assert loss == pytest.approx(loss_b, rel=1e-4)
    assert np.linalg.norm(g - g_b) <= 2e-3 * np.linalg.norm(g_b)
    out = n.eval(x)
    y_b, gu_b = M.input_grad(th, u, bf=True)
    d_b, dd = M.decode(y_b, band, 1.0)
    assert np.median(np.abs(out[:, 0] - d_b)) <= 1e-6
    gx_b = (dd[:, None] * gu_b) * (2.0 / (HI - LO).astype(np.float64))
    gerr = np.linalg.norm(out[:, 1:] - gx_b, axis=1) / (np.linalg.norm(gx_b, axis=1) + 1e-3)
    assert np.median(gerr) <= 1e-4


def test_training_fits_the_dam16k_boundary():
    sc = S.make_scene("dam16k")
    h = float(sc.desc.h)
    steps = 2000
    n = S.NeuralSDF.for_scene(sc.desc, seed=1, band=3 * h, lr_min=1e-5, decay_steps=steps)
    samples = S.sdf_samples(sc.desc, 1 << 19, seed=2)
    n.set_samples(samples)
    ms = n.time_train(10, 65536)
    print(f"training step (batch 65536): {ms / 10:.3f} ms")
    n.set_params(n.get_params())   # restart: fresh Adam state
    losses = n.train(steps, 16384)
    print(f"loss {losses[0]:.3g} -> {losses[-1]:.3g}")
    assert losses[-1] < 0.01 * losses[0]
    # distance error against the FP64 analytic SDF within 2h of the surface
    val = S.sdf_samples(sc.desc, 20000, near_frac=1.0, near_width=2 * sc.desc.h, seed=99)
    pred = n.eval(val[:, :3])
    exact = O.scene_sdf(sc.desc, val[:, :3])
    err = np.abs(pred[:, 0] - exact)
    print(f"|d - d_exact| within 2h: mean {err.mean() / sc.desc.h:.3g} h, p99 {np.quantile(err, 0.99) / sc.desc.h:.3g} h")
    assert err.mean() < 0.5 * sc.desc.h


def test_dam1m_network_against_the_analytic_sdf():
    """The learned SDF of the 1M scene (box + sphere + cone) with the
    training recipe of that workload (4M samples, 70% uniform over the
    boundary grid's box and 30% within 3 h of the surface, batch 65,536,
    20,000 Adam steps, lr 1e-3 with cosine decay), labelled with the exact
    signed distance (FP32 analytic, band-compressed target tanh(d / 3h)):
    its distance error against the FP64 analytic SDF over |d| < r0 (the
    wall-penalty band) within 0.05 h, mean and RMS."""
    sc = S.make_scene("dam1m")
    h = float(sc.desc.h)
    r0 = float(sc.desc.rest_distance)
    steps = 20000
    n = S.NeuralSDF.for_scene(sc.desc, seed=1, band=3 * h, lr=1e-3, lr_min=1e-5, decay_steps=steps)
    n.set_samples(S.sdf_samples(sc.desc, 1 << 22, near_frac=0.3, near_width=3 * h, seed=2))
    losses = n.train(steps, 65536)
    val = S.sdf_samples(sc.desc, 1 << 18, near_frac=1.0, near_width=r0, seed=77)
    exact = O.scene_sdf(sc.desc, val[:, :3])
    band = np.abs(exact) < r0
    pred = n.eval(val[band, :3])[:, 0]
    err = np.abs(pred - exact[band])
    rms = float(np.sqrt((err ** 2).mean()))
    print(f"1M scene, {band.sum()} points with |d| < r0: mean {err.mean() / h:.4f} h, rms {rms / h:.4f} h, "
          f"p99 {np.quantile(err, 0.99) / h:.3f} h, max {err.max() / h:.3f} h; loss {losses[0]:.3g} -> "
          f"{losses[-1]:.3g}")
    assert band.sum() > 100000
    assert err.mean() <= 0.05 * h and rms <= 0.05 * h
