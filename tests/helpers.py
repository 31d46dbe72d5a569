"""Shared test helpers (no product logic)."""
from __future__ import annotations

import numpy as np


def jet_rgb(s: np.ndarray) -> np.ndarray:
    """PortablePixmap::setJET (portable_pixmap.cpp:49-62) in float32, colour
    levels truncated to unsigned as PortablePixmap::set does."""
    s = np.asarray(s, np.float32)
    t = s * np.float32(4.0)
    r = np.minimum(t - np.float32(1.5), -t + np.float32(4.5))
    g = np.minimum(t - np.float32(0.5), -t + np.float32(3.5))
    b = np.minimum(t + np.float32(0.5), -t + np.float32(2.5))
    out = []
    for v in (r, g, b):
        v = np.clip(v, np.float32(0.0), np.float32(1.0)).astype(np.float32)
        out.append((np.float32(255.0) * v).astype(np.uint32))
    return np.stack(out, axis=-1)


def ppm_slice(values_yx: np.ndarray) -> np.ndarray:
    """Image rows as PortablePixmap::set writes them: pixel (i, j) goes to
    row height - j - 1 (portable_pixmap.cpp:44-46)."""
    return jet_rgb(values_yx)[::-1].astype(np.uint8)


def rel_err(a, b, floor):
    return np.abs(np.asarray(a, np.float64) - np.asarray(b, np.float64)) / np.maximum(np.abs(b), floor)


def kinetic_energy(mass: float, vel: np.ndarray) -> float:
    return 0.5 * float(mass) * float((np.asarray(vel, np.float64) ** 2).sum())


def lattice_positions(n: tuple, spacing: float, origin) -> np.ndarray:
    nx, ny, nz = n
    k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    pos = np.stack([origin[0] + i * spacing, origin[1] + j * spacing, origin[2] + k * spacing], -1)
    return pos.reshape(-1, 3)
