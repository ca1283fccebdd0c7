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
