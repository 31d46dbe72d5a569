"""FP64 restatement of the surface-particle classification -- TEST INFRASTRUCTURE.

Only tests/ use this module, as the checker; the product never imports it.

The reference declares extractSurfaceParticles (particle_simulation.h:224)
without defining it; what it does define are the inputs -- cmDistanceThresh
and nPartTresh (h:103-104, 0.5 and 20 in Example01, cu:1811-1812) and the
centre-of-mass sums of compute_particle_acceleration_ifsurf (cu:370-380,
1550-1553: the particle itself plus each neighbour with 0 < r <= h).  The
definition the library implements (include/sph.h, sph_extract_surface):

  n_i = 1 + #{j : 0 < |x_j - x_i| <= h}
  d_i = | sum_j (x_j - x_i) | / n_i
  surface  <=>  n_i < n_thresh  or  d_i > cm_thresh * h

Pinned by closed forms on a cubic lattice of spacing h/2 (tests): the count
of integer points in a ball of radius 2 (33) for an interior particle, the
half-ball count and centre-of-mass offset for a face particle, the
quarter-ball count for an edge particle, and an isolated particle; and by a
brute-force all-pairs evaluation on random points.
"""
from __future__ import annotations

import numpy as np


def surface_flags(pos, h: float, cm_thresh: float, n_thresh: float, brute: bool = False):
    """Flags (uint8, per particle), n_i, d_i, and an 'uncertain' mask: particles
    with a pair within 1e-5 h of the cutoff (a FP32 distance test may count it
    differently) or with d_i within 1e-5 h of the threshold."""
    x = np.asarray(pos, np.float64)
    n = x.shape[0]
    if brute:
        diff = x[None, :, :] - x[:, None, :]
        r = np.sqrt((diff ** 2).sum(-1))
        nb = (r <= h) & (r > 0)
        cnt = 1 + nb.sum(1)
        s = (diff * nb[:, :, None]).sum(1)
        edge = (np.abs(r - h) < 1e-5 * h).any(1)
    else:
        from scipy.spatial import cKDTree
        pairs = cKDTree(x).query_pairs(r=h * (1 + 1e-4), output_type="ndarray")
        i, j = pairs[:, 0], pairs[:, 1]
        dv = x[j] - x[i]
        r = np.sqrt((dv ** 2).sum(1))
        keep = (r <= h) & (r > 0)
        cnt = np.ones(n, np.int64)
        np.add.at(cnt, i[keep], 1)
        np.add.at(cnt, j[keep], 1)
        s = np.zeros((n, 3))
        np.add.at(s, i[keep], dv[keep])
        np.add.at(s, j[keep], -dv[keep])
        near = np.abs(r - h) < 1e-5 * h
        edge = np.zeros(n, bool)
        edge[i[near]] = True
        edge[j[near]] = True
    d = np.sqrt(((s / cnt[:, None]) ** 2).sum(1))
    flags = ((cnt < n_thresh) | (d > cm_thresh * h)).astype(np.uint8)
    uncertain = edge | (np.abs(d - cm_thresh * h) < 1e-5 * h)
    return flags, cnt, d, uncertain
