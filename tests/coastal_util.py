"""Source masks and an independent centroid restatement shared by the
masked-source (coastal) tests."""
import math

import numpy as np


def land_mask(src, seed=5):
    """0 on a 'continent' (lon 20-140, lat -30..60) and a few random cells, -1 on a band
    (lon 190-240, lat > 0) that is wet for the nearest search only, 1 elsewhere."""
    off, idx, c = src["offsets"], src["indices"], src["coords"]
    lon = np.add.reduceat(c[idx, 0], off[:-1]) / np.diff(off)
    lat = np.add.reduceat(c[idx, 1], off[:-1]) / np.diff(off)
    m = np.ones(lon.shape[0], np.int32)
    m[(lon > 20) & (lon < 140) & (lat > -30) & (lat < 60)] = 0
    m[np.random.default_rng(seed).random(lon.shape[0]) < 0.05] = 0
    m[(lon > 190) & (lon < 240) & (lat > 0)] = -1
    return m


def np_centroids(mesh):
    """compute_cell_centroids_xy (weight_generator.cpp:74-138), restated in numpy."""
    off, idx, c = mesh["offsets"], mesh["indices"], mesh["coords"]
    lon = c[idx, 0] * math.pi / 180.0
    lat = c[idx, 1] * math.pi / 180.0
    n = np.diff(off).astype(np.float64)
    x = np.add.reduceat(np.cos(lat) * np.cos(lon), off[:-1]) * (1.0 / n)
    y = np.add.reduceat(np.cos(lat) * np.sin(lon), off[:-1]) * (1.0 / n)
    z = np.add.reduceat(np.sin(lat), off[:-1]) * (1.0 / n)
    la = np.arcsin(z)
    lo = np.arctan2(y, x)
    lo = np.where(lo < 0.0, lo + 2.0 * math.pi, lo)
    return lo * 180.0 / math.pi, la * 180.0 / math.pi
