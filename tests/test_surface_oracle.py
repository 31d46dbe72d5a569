"""Pins the FP64 surface-particle oracle (oracle/surface.py) by closed forms on
a cubic lattice of spacing h/2 and by brute force."""
import numpy as np

from oracle import surface as SF

H = 0.01


def _block(n=9):
    g = np.arange(n) * (H / 2)
    x, y, z = np.meshgrid(g, g, g, indexing="ij")
    return np.stack([x.ravel(), y.ravel(), z.ravel()], 1)


def _index(pos, ijk):
    t = np.asarray(ijk) * (H / 2)
    return int(np.argmin(((pos - t) ** 2).sum(1)))


def test_ball_counts_are_the_closed_forms():
    # integer points (a, b, c) with a^2 + b^2 + c^2 <= 4: 33 in the ball,
    # 23 with c <= 0, 16 with b, c <= 0
    pts = np.array([(a, b, c) for a in range(-2, 3) for b in range(-2, 3) for c in range(-2, 3)
                    if a * a + b * b + c * c <= 4])
    assert len(pts) == 33
    assert (pts[:, 2] <= 0).sum() == 23
    assert ((pts[:, 1] <= 0) & (pts[:, 2] <= 0)).sum() == 16


def test_interior_face_edge_corner_particles():
    # h a hair above two spacings: the shell at exactly 2 spacings is inside
    # whatever the rounding of the lattice coordinates, the next one (sqrt 5)
    # is far outside
    pos = _block()
    h = H * 1.0001
    flags, cnt, d, unc = SF.surface_flags(pos, h, 0.5, 20.0)
    centre = _index(pos, (4, 4, 4))
    face = _index(pos, (4, 4, 8))
    edge = _index(pos, (4, 8, 8))
    corner = _index(pos, (8, 8, 8))
    assert cnt[centre] == 33 and d[centre] < 1e-12 and flags[centre] == 0
    # face: 23 points, centre of mass offset (9 * 1 + 1 * 2) spacings / 23
    assert cnt[face] == 23
    assert abs(d[face] - 11 * (H / 2) / 23) < 1e-12 and flags[face] == 0
    assert cnt[edge] == 16 and flags[edge] == 1        # fewer than 20
    assert flags[corner] == 1
    # a lower cm threshold catches the face particle, not the centre
    f2, _, _, _ = SF.surface_flags(pos, h, 0.2, 20.0)
    assert f2[face] == 1 and f2[centre] == 0
    assert not unc[centre]
    # with h exactly two spacings the outer shell sits on the cutoff: the
    # particles are reported as uncertain
    _, _, _, unc2 = SF.surface_flags(pos, H, 0.5, 20.0)
    assert unc2[centre]


def test_isolated_particle_is_surface():
    pos = np.array([[0.0, 0.0, 0.0], [1.0, 0.0, 0.0]])
    flags, cnt, d, _ = SF.surface_flags(pos, H, 0.5, 20.0)
    assert np.array_equal(cnt, [1, 1]) and np.array_equal(flags, [1, 1]) and np.all(d == 0)


def test_tree_search_matches_brute_force():
    rng = np.random.default_rng(3)
    pos = rng.random((600, 3)) * 0.06
    a = SF.surface_flags(pos, H, 0.3, 12.0)
    b = SF.surface_flags(pos, H, 0.3, 12.0, brute=True)
    for u, v in zip(a, b):
        assert np.allclose(u, v, rtol=0, atol=1e-12)
