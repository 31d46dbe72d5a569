"""Writes tests/golden/meshes.npz from the reference's mesh data files.

The reference ships four triangle meshes (CUDASPHNEW/pyramid.obj,
icosphere.obj, cone.obj, bunny.obj), loaded by TriangleMesh (triangle_mesh.cpp)
for BoundaryMap::AddCanvas (boundary_map.cpp).  They are data; this script
parses their "v" and "f" lines and stores vertices (float32) and 0-based
triangle indices (uint32).  Run where the reference is mounted:

    python tests/golden/make_mesh_fixtures.py /root/reference/CUDASPHNEW
"""
from __future__ import annotations

import os
import sys

import numpy as np


def parse_obj(path: str):
    v, f = [], []
    for line in open(path):
        t = line.split()
        if not t:
            continue
        if t[0] == "v":
            v.append([float(x) for x in t[1:4]])
        elif t[0] == "f":
            idx = [int(x.split("/")[0]) - 1 for x in t[1:]]
            for k in range(1, len(idx) - 1):
                f.append([idx[0], idx[k], idx[k + 1]])
    return np.array(v, np.float32), np.array(f, np.uint32)


def main(ref_dir: str) -> None:
    out = {}
    for name in ("pyramid", "icosphere", "cone", "bunny"):
        v, f = parse_obj(os.path.join(ref_dir, f"{name}.obj"))
        out[f"{name}_vertices"] = v
        out[f"{name}_faces"] = f
        print(f"{name}: {v.shape[0]} vertices, {f.shape[0]} triangles")
    np.savez_compressed(os.path.join(os.path.dirname(os.path.abspath(__file__)), "meshes.npz"), **out)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference/CUDASPHNEW")
