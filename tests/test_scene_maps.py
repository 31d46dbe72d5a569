"""Map solids: obstacles of a scene given by a sampled signed distance (a
mesh's boundary map).  Host-only (scene.cpp), no GPU.

The map here samples the exact distance of a sphere, so the scene grids
built from it must equal those of the analytic sphere obstacle up to the
trilinear error of |x - c| - r (at most 3 dx^2 / (8 (r - band)) ~ 1.6e-4
for the sizes below) and the few ghost-lattice points and grid nodes that
this error moves across a threshold (ghost at depth 0 or support + spacing,
node at |d| = support)."""
import ctypes as C

import numpy as np
import pytest
from scipy import ndimage

from cudasphnew_amd import scenes as S
from cudasphnew_amd._native import SphError, SphMapSolid, SphSolid, as_fptr, lib

# the radius keeps the ghost lattice (which starts at c - r - (support +
# spacing)) off the sphere's exact shells: with r + support + spacing a
# multiple of the spacing, dozens of lattice points would lie exactly at
# depth 0 or support + spacing, where any interpolation error flips them
CENTER, RADIUS, DX = (0.33, 0.21, 0.29), 0.1037, 0.005


def sphere_map(desc, c=CENTER, r=RADIUS, dx=DX, band=None, margin=None, present_all=False):
    if band is None:
        band = desc.boundary_support + S.ghost_spacing_of(desc) + 2 * dx
    if margin is None:
        margin = band + dx
    # extra room below the solid, different per axis: the sphere sits off
    # the grid's centre, so that a transposed node order would move it
    extra = np.array([0.0, 0.013, 0.027])
    lo = (np.array(c) - r - margin - extra).astype(np.float32)
    dims = tuple(int(np.ceil((2 * r + 2 * margin + extra[a]) / dx)) + 1 for a in range(3))
    g = [lo[a] + np.float32(dx) * np.arange(dims[a], dtype=np.float32) for a in range(3)]
    z, y, x = np.meshgrid(g[2], g[1], g[0], indexing="ij")
    d = (np.sqrt((x - c[0]) ** 2.0 + (y - c[1]) ** 2.0 + (z - c[2]) ** 2.0) - r).astype(np.float32)
    p = np.ones_like(d, np.uint8) if present_all else (np.abs(d) <= band).astype(np.uint8)
    return {"dims": dims, "lo": lo, "dx": dx, "distance": d, "present": p, "band": band,
            "mesh_lo": np.array(c, np.float32) - np.float32(r), "mesh_hi": np.array(c, np.float32) + np.float32(r),
            "nodes": np.stack([x, y, z], -1)}


def with_sphere(c=CENTER, r=RADIUS):
    d = S.config_desc("dam16k")
    d.num_obstacles = 1
    d.obstacles[0] = SphSolid(1, (C.c_float * 3)(*c), (C.c_float * 3)(r, 0, 0))
    return d


def test_no_maps_is_the_plain_builder_bit_for_bit():
    desc = with_sphere()
    ref = S.boundary_of(desc)
    nx, ny, nz = (int(v) for v in desc.boundary_dims)
    out = [np.empty((nz, ny, nx), np.float32) for _ in range(3)]
    assert lib.sph_scene_boundary_maps(C.byref(desc), None, 0, *(as_fptr(a) for a in out)) == 0
    for a, b in zip(ref, out):
        assert np.array_equal(a, b)


def test_sampled_sphere_map_gives_the_sphere_grids():
    plain = S.config_desc("dam16k")
    A = S.boundary_of(with_sphere())
    B = S.boundary_of(plain, [sphere_map(plain)])
    dist_a, dist_b = A[0], B[0]
    diff = np.abs(dist_a - dist_b)
    moved = diff > 2e-4
    assert moved.sum() <= 8, moved.sum()
    # the moved nodes sit on the support threshold of the recorded band:
    # one grid holds the node's distance (|d| just under the support, here
    # equal to r0), the other the default -r0
    for g in (dist_a, dist_b):
        assert np.all(np.abs(np.abs(g[moved]) - plain.boundary_support) < 1e-3)
    # density and viscosity: the error against the sphere's own share of
    # the grids (the analytic grids minus those of the empty scene)
    E = S.boundary_of(plain)
    for a, b, e in zip(A[1:], B[1:], E[1:]):
        share = np.abs(a - e)
        assert share.sum() > 0.02 * np.abs(a).sum()
        # a ghost moved across a threshold changes the nodes within one
        # support of it (~30 here): one or two such ghosts, measured 34 and
        # 44 nodes of the sphere's 3926
        off = np.abs(a - b) > 1e-3 * np.abs(a).max()
        assert off.sum() <= 0.03 * (share > 0).sum(), off.sum()
        assert np.abs(a - b).sum() <= 1e-2 * share.sum(), np.abs(a - b).sum() / share.sum()


def test_map_solid_distance_at_nodes_and_sides():
    plain = S.config_desc("dam16k")
    bm = sphere_map(plain)
    band = np.float32(bm["band"])
    # at a node whose neighbours all hold distances the solid's distance is
    # the node's, negated (to the rounding of the node's cell coordinate)
    pres = bm["present"].astype(bool)
    inner = ndimage.binary_erosion(pres, np.ones((3, 3, 3), bool), border_value=0)
    nodes = bm["nodes"][inner][::97]
    want_solid = -bm["distance"][inner][::97]
    assert nodes.shape[0] > 300
    cont = S.scene_sdf(plain, nodes)
    got = S.scene_sdf(plain, nodes, [bm])
    np.testing.assert_allclose(got, np.maximum(cont, want_solid), rtol=0, atol=1e-6)
    # a node at the band's edge has a neighbour without a distance: the
    # cells around it read +-band
    edge = pres & ~inner
    e_nodes = bm["nodes"][edge][::41]
    e_solid = S.scene_sdf(plain, e_nodes, [bm])
    e_cont = S.scene_sdf(plain, e_nodes)
    e_dist = bm["distance"][edge][::41]
    far_side = np.abs(e_dist) > bm["band"] - 2 * DX
    assert far_side.mean() > 0.5
    ok = (e_solid == np.maximum(e_cont, np.float32(bm["band"]) * -np.sign(e_dist))) | \
        (np.abs(e_solid - np.maximum(e_cont, -e_dist)) < 1e-6)
    assert ok.all()
    # deep inside the sphere (no node there holds a distance): +band;
    # inside the grid but far outside the sphere: -band, under the container
    c = np.array(CENTER, np.float32)
    pts = np.array([c, c + [0.02, -0.01, 0.015]], np.float32)
    assert np.all(S.scene_sdf(plain, pts, [bm]) == band)
    far = (bm["lo"] + np.float32(0.002)).reshape(1, 3)
    assert not pres[0, 0, 0]
    assert S.scene_sdf(plain, far, [bm])[0] == max(S.scene_sdf(plain, far)[0], -band)
    # off the grid: outside
    off = np.array([[0.6, 0.6, 0.6]], np.float32)
    assert S.scene_sdf(plain, off, [bm])[0] == S.scene_sdf(plain, off)[0]
    # all nodes present gives the trilinear field everywhere on the grid
    # (at the centre, the apex of |x - c|, within a cell diagonal)
    full = sphere_map(plain, present_all=True)
    assert S.scene_sdf(plain, c.reshape(1, 3), [full])[0] == pytest.approx(RADIUS, abs=np.sqrt(3) * DX / 2)


def test_map_solid_rejects_a_thin_band_or_a_small_grid():
    plain = S.config_desc("dam16k")
    need = plain.boundary_support + S.ghost_spacing_of(plain) + np.sqrt(3) * DX
    with pytest.raises(SphError):
        S.boundary_of(plain, [sphere_map(plain, band=need * 0.95)])
    # the grid must reach support + ghost spacing past the solid's box
    with pytest.raises(SphError):
        S.boundary_of(plain, [sphere_map(plain, margin=0.5 * (plain.boundary_support + S.ghost_spacing_of(plain)))])
    with pytest.raises(SphError):
        S.scene_sdf(plain, np.zeros((1, 3), np.float32), [sphere_map(plain, band=need * 0.5)])
    bad = SphMapSolid()
    assert lib.sph_scene_sdf_maps(C.byref(plain), C.byref(bad), 1, None, 0, None) != 0
    assert lib.sph_scene_sdf_maps(C.byref(plain), None, 1, None, 0, None) != 0


def test_samples_with_maps():
    from cudasphnew_amd import neural as NN
    plain = S.config_desc("dam16k")
    bm = sphere_map(plain)
    # uniform points do not depend on the field: same points, labels of the
    # sphere's field up to the trilinear error
    a = NN.grid_samples(with_sphere(), 20000, near_frac=0.0, seed=5)
    b = NN.grid_samples(plain, 20000, near_frac=0.0, seed=5, maps=[bm])
    np.testing.assert_array_equal(a[:, :3], b[:, :3])
    off = np.abs(a[:, 3] - b[:, 3]) > 2e-4
    assert off.mean() < 1e-3
    assert np.abs(a[:, 3] - b[:, 3]).max() <= plain.rest_distance * 2 + 1e-6
    # the map sphere's band is sampled, labelled with the union's distance
    s = NN.sdf_samples(plain, 5000, near_frac=1.0, near_width=0.01, seed=7, maps=[bm])
    assert np.all(np.abs(s[:, 3]) < 0.01)
    np.testing.assert_array_equal(s[:, 3], S.scene_sdf(plain, s[:, :3], [bm]))
    r = np.linalg.norm(s[:, :3] - np.array(CENTER), axis=1)
    # the sphere's share of the band: its area against the box's, ~5%
    assert (np.abs(r - RADIUS) < 0.011).mean() > 0.03
    # a map the scene cannot use is refused here too
    with pytest.raises(SphError):
        NN.grid_samples(plain, 10, maps=[sphere_map(plain, band=0.01)])
