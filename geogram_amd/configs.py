"""The benchmark workloads, generated deterministically (all draws from
std::mt19937_64 streams, the generator of Numeric::random_float64,
basic/numeric.cpp:76-103, through vgc_sample_points, plus numpy streams for
masses and instance sizes):

* ``cvt_points``                  100k uniform seeds, RNG seed 1 (CVT Lloyd)
* ``mixture_points``              2M seeds, half uniform and half from an
                                  8-component Gaussian mixture, seed 2
* ``uniform_transport_instance``  1M uniform seeds, equal masses, seed 3
* ``gaussian_transport_instance`` 1M seeds of a 16-component mixture,
                                  log-normal masses, seed 4
* ``batched_instances``           256 independent instances, seed counts
                                  log-uniform in [2k, 200k], seed 5
* ``icosphere`` / ``sphere_seeds`` a subdivided icosahedron on the unit
                                  sphere and seeds near it (the surface
                                  restricted Voronoi diagram)
"""
from __future__ import annotations

import numpy as np


def cvt_points(n: int = 100_000):
    from . import sample_points
    return sample_points(0, 1, n)


def mixture_points(n: int = 2_000_000):
    from . import sample_points
    return sample_points(1, 2, n)


def uniform_transport_instance(n: int = 1_000_000):
    from . import sample_points
    return sample_points(0, 3, n), np.full(n, 1.0 / n)


def gaussian_transport_instance(n: int = 1_000_000):
    from .warmstart import make_instance
    return make_instance(4, n)


def batched_instances(count: int = 256, nmin: int = 2000, nmax: int = 200_000):
    """Instance k: uniform seeds for even k, 16-Gaussian seeds for odd k;
    equal masses for k % 4 < 2, log-normal otherwise; unit-cube boxes.
    Returns (points (N, 3), masses (N,), sizes (count,), boxes (count, 6))."""
    from . import sample_points
    rng = np.random.default_rng(5)
    sizes = np.exp(rng.uniform(np.log(nmin), np.log(nmax), count)).astype(np.int64)
    xs, nus = [], []
    for k, sz in enumerate(sizes):
        xs.append(sample_points(0 if k % 2 == 0 else 2, 5000 + k, int(sz)))
        if k % 4 < 2:
            nuk = np.full(int(sz), 1.0 / sz)
        else:
            nuk = np.exp(np.random.default_rng(k).normal(0.0, 0.5, int(sz)))
            nuk /= nuk.sum()
        nus.append(nuk)
    boxes = np.array([[0, 0, 0, 1, 1, 1]] * count, dtype=float)
    return np.vstack(xs), np.concatenate(nus), sizes, boxes


def icosphere(level: int, radius: float = 1.0, center=(0.0, 0.0, 0.0)):
    """Triangulated sphere: the icosahedron with each triangle split in four
    level times, new vertices pushed onto the sphere.  Returns (V (nv, 3)
    float64, F (20 * 4**level, 3) int32), triangles counter-clockwise seen
    from outside."""
    t = (1.0 + 5.0 ** 0.5) / 2.0
    v = [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t), (0, -1, -t),
         (0, 1, -t), (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)]
    f = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4),
         (11, 10, 2), (10, 7, 6), (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9),
         (4, 9, 5), (2, 4, 11), (6, 2, 10), (8, 6, 7), (9, 8, 1)]
    V = [np.array(p, dtype=np.float64) / np.linalg.norm(p) for p in v]
    F = np.array(f, dtype=np.int64)
    for _ in range(level):
        edges = np.sort(np.concatenate([F[:, [0, 1]], F[:, [1, 2]], F[:, [2, 0]]]), axis=1)
        uniq, inv = np.unique(edges, axis=0, return_inverse=True)
        base = len(V)
        mids = np.asarray(V)[uniq[:, 0]] + np.asarray(V)[uniq[:, 1]]
        mids /= np.linalg.norm(mids, axis=1)[:, None]
        V = list(np.asarray(V)) + list(mids)
        nf = F.shape[0]
        ab, bc, ca = (base + inv.reshape(-1)[:nf], base + inv.reshape(-1)[nf:2 * nf],
                      base + inv.reshape(-1)[2 * nf:])
        a, b, c = F[:, 0], F[:, 1], F[:, 2]
        F = np.concatenate([np.stack([a, ab, ca], 1), np.stack([b, bc, ab], 1),
                            np.stack([c, ca, bc], 1), np.stack([ab, bc, ca], 1)])
    V = np.asarray(V) * radius + np.asarray(center, dtype=np.float64)
    return V, F.astype(np.int32)


def sphere_seeds(n: int, seed: int, radius: float = 1.0, jitter: float = 0.01):
    """n seeds near the unit sphere (uniform directions, radius jittered by a
    relative normal of scale jitter)."""
    rng = np.random.default_rng(seed)
    x = rng.standard_normal((n, 3))
    x /= np.linalg.norm(x, axis=1)[:, None]
    return x * radius * (1.0 + jitter * rng.standard_normal((n, 1)))
