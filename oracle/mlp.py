"""FP64 restatement of the neural boundary SDF -- TEST INFRASTRUCTURE.

Only tests/, __graft_entry__.smoke() and the cpu_baseline leg of bench.py
use this module, as the checker; the product never imports it.

The network has no reference counterpart (the reference samples its
boundary from dense 3D textures, particle_simulation.cu:48-53, 2959-3011);
it is the learned stand-in for that grid.  Definition:

  u   = 2 (x - lo) / (hi - lo) - 1                      (per axis, FP32)
  g   = [u_x, u_y, u_z,
         sin(pi f_k u_a), cos(pi f_k u_a)  for k < 16, a < 3]   (99 features,
         index 3 + 6k + 2a + {0: sin, 1: cos}, zero-padded to 128)
  f_k = 2^(7k/15) (rounded to FP32)
  z1 = g W1 + b1,  a1 = silu(z1)       (128 -> 64)
  z2 = a1 W2 + b2, a2 = silu(z2)       (64 -> 64)
  z3 = a2 W3 + b3, a3 = silu(z3)       (64 -> 64)
  y  = a3 . w4 + b4,   d = out_scale * y

Target and decoding: t(d) = tanh(d / band), d = band * atanh(y) (band > 0),
else t(d) = d / out_scale, d = out_scale * y.
Loss over a batch of B samples: L = (1/B) sum (y - t(d_true))^2.

With bf16=True the operands of the matrix products are rounded to bfloat16
(round to nearest even) where the device kernels round them: the features,
W1..W3, a1, a2 and the back-propagated dz1..dz3; biases, the output layer,
products and sums stay at full precision.  That models the device's BF16
MFMA inputs with FP32 accumulation.

Pinned by: the finite-difference test of the FP64 gradient (tests), the
closed form of the features, and Adam's published update rule (Kingma & Ba,
2015, Algorithm 1).
"""
from __future__ import annotations

import numpy as np

NF = 16
NFEAT = 3 + 6 * NF        # 99
K1 = 128
H = 64
# f_k = 2^(7k/15), rounded to FP32 as the device holds them
FREQS = (2.0 ** (7.0 * np.arange(NF) / 15.0)).astype(np.float32).astype(np.float64)

# flat parameter layout (FP32 master copy on the device)
O_W1 = 0
O_B1 = O_W1 + K1 * H
O_W2 = O_B1 + H
O_B2 = O_W2 + H * H
O_W3 = O_B2 + H
O_B3 = O_W3 + H * H
O_W4 = O_B3 + H
O_B4 = O_W4 + H
NUM_PARAMS = O_B4 + 1


def unpack(theta: np.ndarray) -> dict:
    t = np.asarray(theta, np.float64)
    return {
        "W1": t[O_W1:O_B1].reshape(K1, H), "b1": t[O_B1:O_W2],
        "W2": t[O_W2:O_B2].reshape(H, H), "b2": t[O_B2:O_W3],
        "W3": t[O_W3:O_B3].reshape(H, H), "b3": t[O_B3:O_W4],
        "w4": t[O_W4:O_B4], "b4": t[O_B4:O_B4 + 1],
    }


def pack(p: dict) -> np.ndarray:
    return np.concatenate([p["W1"].ravel(), p["b1"], p["W2"].ravel(), p["b2"], p["W3"].ravel(),
                           p["b3"], p["w4"], p["b4"]]).astype(np.float64)


def bf16(x: np.ndarray) -> np.ndarray:
    """Round to bfloat16 (nearest even) through FP32, returned as float64."""
    f = np.ascontiguousarray(x, np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def normalized(x: np.ndarray, lo, hi) -> np.ndarray:
    """u in FP32 arithmetic: 2 * ((x - lo) / (hi - lo)) - 1."""
    x = np.asarray(x, np.float32)
    lo = np.asarray(lo, np.float32)
    ext = (np.asarray(hi, np.float32) - lo).astype(np.float32)
    t = ((x - lo).astype(np.float32) / ext).astype(np.float32)
    return (np.float32(2.0) * t - np.float32(1.0)).astype(np.float32)


def features(u: np.ndarray) -> np.ndarray:
    u = np.asarray(u, np.float64)
    n = u.shape[0]
    g = np.zeros((n, K1))
    g[:, :3] = u
    for k in range(NF):
        arg = np.pi * FREQS[k] * u            # (n, 3)
        g[:, 3 + 6 * k + 0:3 + 6 * k + 6:2] = np.sin(arg)
        g[:, 3 + 6 * k + 1:3 + 6 * k + 6:2] = np.cos(arg)
    return g


def feature_du(u: np.ndarray) -> np.ndarray:
    """d g_f / d u_a as (n, K1, 3)."""
    u = np.asarray(u, np.float64)
    n = u.shape[0]
    j = np.zeros((n, K1, 3))
    for a in range(3):
        j[:, a, a] = 1.0
    for k in range(NF):
        w = np.pi * FREQS[k]
        for a in range(3):
            j[:, 3 + 6 * k + 2 * a, a] = w * np.cos(w * u[:, a])
            j[:, 3 + 6 * k + 2 * a + 1, a] = -w * np.sin(w * u[:, a])
    return j


def target(d, band: float, out_scale: float):
    d = np.asarray(d, np.float64)
    return np.tanh(d / band) if band > 0 else d / out_scale


def decode(y, band: float, out_scale: float):
    """Distance and d(distance)/dy."""
    y = np.asarray(y, np.float64)
    if band > 0:
        yc = np.clip(y, -0.999999, 0.999999)
        return band * np.arctanh(yc), band / (1.0 - yc * yc)
    return out_scale * y, np.full_like(y, out_scale)


def silu(z):
    return z / (1.0 + np.exp(-z))


def dsilu(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def forward(theta, u, bf: bool = False):
    p = unpack(theta)
    r = bf16 if bf else (lambda a: np.asarray(a, np.float64))
    g = r(features(u))
    z1 = g @ r(p["W1"]) + p["b1"]
    a1 = silu(z1)
    z2 = r(a1) @ r(p["W2"]) + p["b2"]
    a2 = silu(z2)
    z3 = r(a2) @ r(p["W3"]) + p["b3"]
    a3 = silu(z3)
    y = a3 @ p["w4"] + p["b4"][0]
    return y, dict(g=g, z1=z1, a1=a1, z2=z2, a2=a2, z3=z3, a3=a3)


def loss_and_grad(theta, u, y_true, bf: bool = False):
    """Loss and flat gradient for one batch (u: normalized coordinates,
    y_true = d_true / out_scale)."""
    p = unpack(theta)
    r = bf16 if bf else (lambda a: np.asarray(a, np.float64))
    y, c = forward(theta, u, bf)
    B = y.shape[0]
    e = y - y_true
    loss = float(np.mean(e * e))
    gy = 2.0 * e / B
    grads = {"w4": c["a3"].T @ gy, "b4": np.array([gy.sum()])}
    dz3 = np.outer(gy, p["w4"]) * dsilu(c["z3"])
    grads["b3"] = dz3.sum(0)
    grads["W3"] = r(c["a2"]).T @ r(dz3)
    dz2 = (r(dz3) @ r(p["W3"]).T) * dsilu(c["z2"])
    grads["b2"] = dz2.sum(0)
    grads["W2"] = r(c["a1"]).T @ r(dz2)
    dz1 = (r(dz2) @ r(p["W2"]).T) * dsilu(c["z1"])
    grads["b1"] = dz1.sum(0)
    grads["W1"] = c["g"].T @ r(dz1)
    return loss, pack(grads)


def input_grad(theta, u, bf: bool = False):
    """y and dy/du (n, 3)."""
    p = unpack(theta)
    r = bf16 if bf else (lambda a: np.asarray(a, np.float64))
    y, c = forward(theta, u, bf)
    dz3 = np.outer(np.ones_like(y), p["w4"]) * dsilu(c["z3"])
    dz2 = (r(dz3) @ r(p["W3"]).T) * dsilu(c["z2"])
    dz1 = (r(dz2) @ r(p["W2"]).T) * dsilu(c["z1"])
    dg = r(dz1) @ r(p["W1"]).T
    return y, np.einsum("nf,nfa->na", dg, feature_du(u))


def adam(theta, m, v, grad, t, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam update (Kingma & Ba 2015, Algorithm 1), t = 1-based step."""
    m = b1 * m + (1.0 - b1) * grad
    v = b2 * v + (1.0 - b2) * grad * grad
    mh = m / (1.0 - b1 ** t)
    vh = v / (1.0 - b2 ** t)
    return theta - lr * mh / (np.sqrt(vh) + eps), m, v


def xavier_init(seed: int) -> np.ndarray:
    """Not the device's initializer (that one is splitmix64-based); used only
    to build random networks for the checks."""
    rng = np.random.default_rng(seed)
    p = {}
    for name, (fi, fo) in {"W1": (NFEAT, H), "W2": (H, H), "W3": (H, H)}.items():
        lim = np.sqrt(6.0 / (fi + fo))
        w = np.zeros((K1 if name == "W1" else H, H))
        w[:fi] = rng.uniform(-lim, lim, (fi, fo))
        p[name] = w
    p["w4"] = rng.uniform(-np.sqrt(6.0 / (H + 1)), np.sqrt(6.0 / (H + 1)), H)
    for b in ("b1", "b2", "b3"):
        p[b] = rng.uniform(-0.1, 0.1, H)
    p["b4"] = np.array([0.05])
    return pack(p)
