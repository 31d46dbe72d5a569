"""Mesh signed distance field generated on the GPU against the FP64 oracle.

Bars: node set (distance <= maxDist) identical except nodes within 1e-5 of
maxDist; distance within 1e-5 absolute (FP32 vs FP64 over coordinates of
size ~1: the Voronoi-region products of the closest-point test cancel near
sharp vertices, e.g. the cone's tip, measured up to ~25 ulp); sign identical where |distance| > 1e-5 (same parity rule, exact
orientation signs on both sides); unit normals within 1e-3 except where two
triangles' closest points are equally near (medial nodes, <= 0.5%)."""
import os

import numpy as np
import pytest

from cudasphnew_amd import meshsdf as M
from oracle import oracle as O

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
MESHES = np.load(os.path.join(HERE, "golden", "meshes.npz"))


def _compare(v, f, rest, cs, dx):
    cfg = M.config(rest, cs, dx)
    bm = M.generate(v, f, cfg)
    d, n, p = O.mesh_sdf(v, f, rest, cs, dx, bm["dims"], bm["lo"], bm["hi"])
    maxd = float(np.float32(max(rest, cs)))
    edge = np.abs(np.abs(d) - maxd) < 1e-5
    assert np.array_equal(bm["present"][~edge], p[~edge])
    both = (bm["present"] == 1) & (p == 1)
    err = np.abs(bm["distance"][both] - d[both])
    assert err.max() <= 1e-5, err.max()
    clear = both & (np.abs(d) > 1e-5)
    assert np.array_equal(bm["distance"][clear] < 0, d[clear] < 0)
    nz = both & (np.abs(d) > 1e-5)
    dots = np.sum(bm["normal"][nz] * n[nz], -1)
    bad = dots < 1 - 1e-3
    assert bad.mean() <= 5e-3, bad.mean()
    assert np.all(bm["distance"][bm["present"] == 0] == np.float32(rest))
    return bm, int(both.sum()), float((d[both] < 0).mean())


@pytest.mark.parametrize("name", ["pyramid", "icosphere", "cone", "bunny"])
def test_mesh_sdf_matches_oracle(name):
    v = M.fit(MESHES[f"{name}_vertices"], (-0.75, -0.75, -0.75), (0.75, 0.75, 0.75))
    f = MESHES[f"{name}_faces"]
    cs = 0.05
    bm, npresent, frac_inside = _compare(v, f, 2 * cs, cs, cs / 3)
    print(f"{name}: dims {bm['dims']}, {npresent} nodes within maxDist, {frac_inside:.2f} inside, "
          f"{bm['ms']:.3f} ms")
    assert npresent > 1000 and 0.2 < frac_inside < 0.8


def test_reference_example_configuration():
    """main.cpp:91-103: icosphere fitted to [-0.75, 0.75]^3, compactSupport
    0.017, dx = compactSupport / 3, restDistance = 2 compactSupport."""
    v = M.fit(MESHES["icosphere_vertices"], (-0.75, -0.75, -0.75), (0.75, 0.75, 0.75))
    f = MESHES["icosphere_faces"]
    cs = np.float32(0.017)
    bm, npresent, frac_inside = _compare(v, f, float(np.float32(cs * np.float32(2))), float(cs),
                                         float(cs / np.float32(3)))
    print(f"icosphere 0.017: dims {bm['dims']}, {npresent} nodes, {bm['ms']:.2f} ms")
    assert bm["dims"][0] > 250


def test_generated_map_save_load(tmp_path):
    v = M.fit(MESHES["cone_vertices"], (-0.5, -0.5, -0.5), (0.5, 0.5, 0.5))
    f = MESHES["cone_faces"]
    cfg = M.config(0.06, 0.03, 0.01)
    bm = M.generate(v, f, cfg)
    M.save(str(tmp_path / "cone.txt"), cfg, bm)
    cfg2, bm2 = M.load(str(tmp_path / "cone.txt"))
    assert np.array_equal(bm2["present"], bm["present"])
    assert np.allclose(bm2["distance"], bm["distance"], rtol=1e-7, atol=0)
    assert np.allclose(bm2["normal"], bm["normal"], rtol=1e-7, atol=0)


def test_network_fits_a_generated_mesh_sdf():
    """The learned SDF trained on the band nodes of a generated boundary map
    (the reference's example configuration, icosphere, 266^3 nodes): RMS
    error over held-out band nodes within 0.05 compactSupport, and the sign
    right at 98% of the held-out nodes within compactSupport of the surface
    (the ones it misses lie within the network's largest error, ~0.3 cs, of
    the surface)."""
    from cudasphnew_amd import neural as N
    v = M.fit(MESHES["icosphere_vertices"], (-0.75, -0.75, -0.75), (0.75, 0.75, 0.75))
    f = MESHES["icosphere_faces"]
    cs = 0.017
    bm = M.generate(v, f, M.config(2 * cs, cs, cs / 3))
    smp = N.boundary_map_samples(bm)
    rng = np.random.default_rng(1)
    perm = rng.permutation(smp.shape[0])
    test, train = smp[perm[: smp.shape[0] // 10]], smp[perm[smp.shape[0] // 10:]]
    net = N.train_for_boundary_map(bm, band=2 * cs, samples=train)
    pred = net.eval(test[:, :3])[:, 0]
    err = np.abs(pred - test[:, 3])
    near = np.abs(test[:, 3]) < cs
    rms = float(np.sqrt((err ** 2).mean()))
    sign = float((np.sign(pred[near]) == np.sign(test[near, 3])).mean())
    print(f"icosphere band nodes {smp.shape[0]}: rms {rms / cs:.4f} cs, p99 {np.quantile(err, 0.99) / cs:.3f} cs, "
          f"sign agreement {sign:.4f}")
    assert rms <= 0.05 * cs
    assert sign >= 0.98
    wrong = near.copy()
    wrong[near] = np.sign(pred[near]) != np.sign(test[near, 3])
    assert np.all(np.abs(test[wrong, 3]) <= 0.35 * cs)
