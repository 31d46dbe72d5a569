"""Writes tests/golden/ppm_fixtures.npz from the reference's golden slices.

The reference ships two PPM slices of its boundary grids that match its
current code (SphInComplexShapes::SaveSlicedDistanceMapToPpm /
SaveSlicedViscosityMapToPpm, called from particle_simulation.cu:2613-2614):
    CUDASPHNEW/distances.ppm    (259 x 194, P3)
    CUDASPHNEW/viscosities.ppm  (259 x 194, P3)
They are data files of the reference; this script only decodes the ASCII P3
format and stores the RGB triples as uint8 arrays (row 0 = first image row,
as written).  Run it in a container where the reference is mounted:

    python tests/golden/make_ppm_fixtures.py /root/reference/CUDASPHNEW
"""
from __future__ import annotations

import hashlib
import os
import sys

import numpy as np


def read_p3(path: str) -> np.ndarray:
    with open(path, "r") as f:
        tokens = f.read().split()
    if tokens[0] != "P3":
        raise ValueError(f"{path}: not an ASCII P3 image")
    w, h, maxc = int(tokens[1]), int(tokens[2]), int(tokens[3])
    vals = np.array([int(t) for t in tokens[4:4 + 3 * w * h]], dtype=np.int64)
    if vals.size != 3 * w * h or maxc != 255:
        raise ValueError(f"{path}: unexpected size or max colour")
    return vals.reshape(h, w, 3).astype(np.uint8)


def main(ref_dir: str) -> None:
    out = {}
    for name in ("distances", "viscosities"):
        img = read_p3(os.path.join(ref_dir, f"{name}.ppm"))
        out[name] = img
        out[name + "_sha256"] = np.frombuffer(hashlib.sha256(img.tobytes()).digest(), np.uint8)
        print(name, img.shape, hashlib.sha256(img.tobytes()).hexdigest())
    dst = os.path.join(os.path.dirname(os.path.abspath(__file__)), "ppm_fixtures.npz")
    np.savez_compressed(dst, **out)
    print("wrote", dst, os.path.getsize(dst), "bytes")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference/CUDASPHNEW")
