"""Writes tests/golden/ref_geometry_vectors.npz: inputs and outputs of the
reference's own sphere geometry (libs/axis/include/axis/detail/
spherical_geometry.hpp, compiled from the reference tree into oracle/_ref by
`make -C oracle ref`), so that the oracle's pins against it hold where the
reference tree is absent.

    python tests/golden/make_ref_vectors.py

Vectors: lonlat_to_xyz at 2,000 random points and at the nodes of the 2 and
0.25 degree lat-lon grids' first rows (radians: deg * (pi / 180), the
weight generator's conversion); the polygon area of 300 cap-inscribed CCW
polygons (3 to 11 vertices, radii 1e-4 to 0.5 rad).
"""
import ctypes as C
import math
import pathlib
import sys

import numpy as np

HERE = pathlib.Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parent.parent / "oracle"))
import oracle_py as O  # noqa: E402


def cap_polygon(lon, lat, radius, n):
    """CCW n-gon inscribed in the cap of angular radius `radius` (radians)."""
    lo, la = math.radians(lon), math.radians(lat)
    c = np.array([math.cos(la) * math.cos(lo), math.cos(la) * math.sin(lo), math.sin(la)])
    e = np.array([-math.sin(lo), math.cos(lo), 0.0])
    n_ = np.cross(c, e)
    t = np.linspace(0, 2 * math.pi, n, endpoint=False)
    p = math.cos(radius) * c[None, :] + math.sin(radius) * (np.cos(t)[:, None] * e[None, :] + np.sin(t)[:, None] * n_[None, :])
    return p / np.linalg.norm(p, axis=1)[:, None]


def main():
    R = O.ref_lib()
    if R is None:
        raise SystemExit("oracle/_ref/libref_geometry.so is not built (make -C oracle ref, with the reference tree present)")
    D = C.POINTER(C.c_double)
    rng = np.random.default_rng(3)
    pts = [(rng.uniform(-7, 7), rng.uniform(-1.6, 1.6)) for _ in range(2000)]
    d2r = math.pi / 180.0
    pts += [(i * 2.0 * d2r, -90.0 * d2r) for i in range(181)] + [(i * 0.25 * d2r, -89.75 * d2r) for i in range(1441)]
    ll = np.array(pts)
    xyz = np.zeros((len(pts), 3))
    out = np.zeros(3)
    for k, (lo, la) in enumerate(pts):
        R.ref_lonlat_to_xyz(lo, la, out.ctypes.data_as(D))
        xyz[k] = out
    rng = np.random.default_rng(11)
    polys, counts, areas = [], [], []
    for _ in range(300):
        a = np.ascontiguousarray(cap_polygon(rng.uniform(0, 360), rng.uniform(-85, 85), rng.uniform(1e-4, 0.5),
                                             int(rng.integers(3, 12))))
        polys.append(a)
        counts.append(a.shape[0])
        areas.append(R.ref_polygon_area(a.ctypes.data_as(D), a.shape[0]))
    np.savez_compressed(HERE / "ref_geometry_vectors.npz", lonlat=ll, xyz=xyz, poly_xyz=np.concatenate(polys),
                        poly_counts=np.array(counts), poly_area=np.array(areas))
    print("wrote", HERE / "ref_geometry_vectors.npz", len(pts), "points", len(polys), "polygons")


if __name__ == "__main__":
    main()
