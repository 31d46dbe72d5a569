"""The C ABI from plain C: a C99 program includes include/sph.h, links
libsph_hip.so and drives the host-only part of the boundary (version,
device count, the dam16k scene description, its particles and boundary
grids), as a binding written against the header would.  No GPU needed; the
grids are compared with the Python wrapper's."""
import os
import shutil
import subprocess

import numpy as np
import pytest

from cudasphnew_amd import _native as N
from cudasphnew_amd import scenes as S

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

PROGRAM = r"""
#include <stdio.h>
#include <stdlib.h>
#include "sph.h"

int main(int argc, char** argv) {
    if (argc < 2) return 2;
    if (sph_abi_version() != SPH_ABI_VERSION) return 3;
    printf("devices %d\n", (int)sph_device_count());
    sph_scene_desc d;
    if (sph_scene_config(SPH_CONFIG_DAM16K, 0, 1, &d) != SPH_OK) return 4;   /* seed 1, as make_scene */
    sph_params p;
    if (sph_scene_params(&d, &p) != SPH_OK) return 5;
    uint32_t n = 0;
    if (sph_scene_particles(&d, NULL, 0, &n) != SPH_OK) return 6;
    float* pos = (float*)malloc(sizeof(float) * 3 * n);
    if (sph_scene_particles(&d, pos, n, &n) != SPH_OK) return 7;
    size_t nodes = (size_t)d.boundary_dims[0] * d.boundary_dims[1] * d.boundary_dims[2];
    float* g = (float*)malloc(sizeof(float) * 3 * nodes);
    if (sph_scene_boundary(&d, g, g + nodes, g + 2 * nodes) != SPH_OK) return 8;
    /* an invalid call reports an error instead of failing silently */
    if (sph_scene_boundary(NULL, g, g, g) != SPH_INVALID_ARG) return 9;
    printf("particles %u cells %d %d %d h %.6f\n", n, p.grid_dim[0], p.grid_dim[1], p.grid_dim[2],
           (double)p.compact_support);
    FILE* f = fopen(argv[1], "wb");
    if (!f) return 10;
    fwrite(pos, sizeof(float), 3 * (size_t)n, f);
    fwrite(g, sizeof(float), 3 * nodes, f);
    fclose(f);
    free(pos);
    free(g);
    return 0;
}
"""


@pytest.mark.skipif(shutil.which("gcc") is None, reason="needs gcc")
def test_plain_c_program_uses_the_abi(tmp_path):
    src = tmp_path / "client.c"
    src.write_text(PROGRAM)
    exe = tmp_path / "client"
    libdir = os.path.dirname(N.LIB_PATH)
    subprocess.run(["gcc", "-std=c99", "-Wall", "-Wextra", "-pedantic", "-Werror", f"-I{ROOT}/include", str(src),
                    "-o", str(exe), f"-L{libdir}", "-lsph_hip", f"-Wl,-rpath,{libdir}"], check=True)
    out = tmp_path / "out.bin"
    r = subprocess.run([str(exe), str(out)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, (r.returncode, r.stdout, r.stderr)
    sc = S.make_scene("dam16k")
    assert f"particles {sc.n} " in r.stdout
    data = np.fromfile(out, np.float32)
    nodes = int(np.prod(sc.boundary_dims))
    assert data.size == 3 * sc.n + 3 * nodes
    np.testing.assert_array_equal(data[:3 * sc.n].reshape(-1, 3), sc.positions)
    grids = data[3 * sc.n:].reshape(3, -1)
    for k, name in enumerate(("distance", "density", "viscosity")):
        np.testing.assert_array_equal(grids[k], getattr(sc, name).ravel())
