"""Mesh generators in the AXIS UnstructuredMesh layout.

A mesh is a dict with
  coords   float64 [n_nodes, 2] in column-major (layout_left) order: the
           array handed to the library is ``coords.T.copy()`` flattened, i.e.
           all longitudes, then all latitudes (degrees);
  offsets  int64 [n_cells + 1]  CSR cell -> node offsets;
  indices  int64 [offsets[-1]]  node indices, counter-clockwise seen from
           outside the sphere (a clockwise clip polygon gets zero coverage in
           the great-circle clipper, libs/axis/benchmarks/README.md:94).

regular_latlon follows RuleGenerator's RegularLatLon (reference
libs/axis/src/topology/rule_generator.cpp:105-169): nodes at min + i*r,
node (i, j) at j*(ni+1) + i, cell (i, j) at i + j*ni with corners BL, BR, TR,
TL.  The reference ships no cubed-sphere or Voronoi generator, so those are
this library's own.
"""
from __future__ import annotations

import math

import numpy as np


def _mesh(lon, lat, offsets, indices):
    coords = np.empty((lon.shape[0], 2), dtype=np.float64, order="F")
    coords[:, 0] = lon
    coords[:, 1] = lat
    return {
        "coords": coords,
        "offsets": np.ascontiguousarray(offsets, dtype=np.int64),
        "indices": np.ascontiguousarray(indices, dtype=np.int64),
    }


def n_cells(mesh) -> int:
    return int(mesh["offsets"].shape[0] - 1)


def n_nodes(mesh) -> int:
    return int(mesh["coords"].shape[0])


def regular_latlon(ni: int, nj: int, min_x: float = 0.0, min_y: float = -90.0, r_x: float | None = None,
                   r_y: float | None = None):
    """Regular lat-lon grid of ni x nj cells (RuleGenerator semantics)."""
    if r_x is None:
        r_x = 360.0 / ni
    if r_y is None:
        r_y = 180.0 / nj
    i = np.arange(ni + 1, dtype=np.float64)
    j = np.arange(nj + 1, dtype=np.float64)
    lon_row = min_x + i * r_x
    lat_col = min_y + j * r_y
    lon = np.tile(lon_row, nj + 1)
    lat = np.repeat(lat_col, ni + 1)
    ci = np.tile(np.arange(ni, dtype=np.int64), nj)
    cj = np.repeat(np.arange(nj, dtype=np.int64), ni)
    stride = ni + 1
    idx = np.stack([cj * stride + ci, cj * stride + ci + 1, (cj + 1) * stride + ci + 1, (cj + 1) * stride + ci], axis=1)
    offsets = np.arange(0, 4 * ni * nj + 1, 4, dtype=np.int64)
    return _mesh(lon, lat, offsets, idx.reshape(-1))


def latlon_deg(res_deg: float):
    """Global lat-lon grid with edges at multiples of res_deg, lon in [0, 360]."""
    ni = int(round(360.0 / res_deg))
    nj = int(round(180.0 / res_deg))
    return regular_latlon(ni, nj, 0.0, -90.0, res_deg, res_deg)


# ----------------------------------------------------------------- cubed sphere
_FACES = (
    # (fixed axis, fixed side, axis of local i, sign, axis of local j, sign)
    # equatorial faces, i grows eastward, j northward
    (0, +1, 1, +1, 2, +1),   # +X
    (1, +1, 0, -1, 2, +1),   # +Y
    (0, -1, 1, -1, 2, +1),   # -X
    (1, -1, 0, +1, 2, +1),   # -Y
    (2, +1, 1, +1, 0, -1),   # +Z (north)
    (2, -1, 1, +1, 0, +1),   # -Z (south)
)


def cubed_sphere(n: int):
    """Gnomonic equiangular cubed sphere C<n>: 6 n^2 quads, 6 n^2 + 2 welded nodes.

    Nodes sit on an integer lattice of the cube surface, so nodes shared by
    two or three faces weld exactly.  Their positions are tan-spaced
    (equiangular), projected to the unit sphere and stored as lon in [0, 360),
    lat in [-90, 90] degrees.  Cell index = face * n^2 + j * n + i.
    """
    half = n // 2
    t = np.empty(n + 1, dtype=np.float64)
    for k in range(n + 1):
        if 2 * k < n:
            t[k] = math.tan(-math.pi / 4 + k * (math.pi / 2) / n)
        elif 2 * k == n:
            t[k] = 0.0
        else:
            t[k] = -t[n - k]
    t[0], t[n] = -1.0, 1.0
    del half

    m = n + 1
    a = np.arange(m, dtype=np.int64)[None, :]
    b = np.arange(m, dtype=np.int64)[:, None]
    keys = []
    for (fa, fs, ia, isg, ja, jsg) in _FACES:
        L = [None, None, None]
        L[fa] = np.full((m, m), n if fs > 0 else 0, dtype=np.int64)
        L[ia] = np.broadcast_to(a if isg > 0 else n - a, (m, m))
        L[ja] = np.broadcast_to(b if jsg > 0 else n - b, (m, m))
        keys.append((L[0] * m + L[1]) * m + L[2])
    keys = np.stack(keys)  # [6, n+1, n+1] lattice key of node (face, b, a)
    uniq, first, inv = np.unique(keys.reshape(-1), return_index=True, return_inverse=True)
    # number nodes in order of first appearance (face-major, row-major)
    order = np.argsort(first, kind="stable")
    rank = np.empty_like(order)
    rank[order] = np.arange(order.shape[0])
    grid = rank[inv.reshape(-1)].reshape(6, m, m)
    cells = np.empty((6, n, n, 4), dtype=np.int64)
    cells[:, :, :, 0] = grid[:, :-1, :-1]
    cells[:, :, :, 1] = grid[:, :-1, 1:]
    cells[:, :, :, 2] = grid[:, 1:, 1:]
    cells[:, :, :, 3] = grid[:, 1:, :-1]
    uk = uniq[order]
    L = np.stack([uk // (m * m), (uk // m) % m, uk % m], axis=1)
    xyz = t[L]
    xyz /= np.sqrt((xyz * xyz).sum(axis=1))[:, None]
    lon = np.degrees(np.arctan2(xyz[:, 1], xyz[:, 0]))
    lon = np.where(lon < 0.0, lon + 360.0, lon)
    lon = np.where(lon >= 360.0, lon - 360.0, lon)
    lat = np.degrees(np.arctan2(xyz[:, 2], np.hypot(xyz[:, 0], xyz[:, 1])))
    idx = cells.reshape(-1, 4)
    idx = _orient_ccw(idx, lon, lat)
    offsets = np.arange(0, 4 * idx.shape[0] + 1, 4, dtype=np.int64)
    return _mesh(lon, lat, offsets, idx.reshape(-1))


def _xyz_deg(lon, lat):
    lo, la = np.radians(lon), np.radians(lat)
    return np.stack([np.cos(la) * np.cos(lo), np.cos(la) * np.sin(lo), np.sin(la)], axis=-1)


def _orient_ccw(idx, lon, lat):
    """Reverse any quad whose vertex ring is clockwise seen from outside."""
    p = _xyz_deg(lon, lat)[idx]  # [nc, 4, 3]
    c = p.mean(axis=1)
    s = np.zeros(idx.shape[0])
    for k in range(idx.shape[1]):
        a, b = p[:, k], p[:, (k + 1) % idx.shape[1]]
        s += np.einsum("ij,ij->i", np.cross(a, b), c)
    flip = s < 0
    out = idx.copy()
    out[flip] = out[flip][:, ::-1]
    return out


# ------------------------------------------------------------------- Voronoi
def icosahedron_vertices(level: int):
    """Vertices of a level-`level` subdivided icosahedron (10*4^level + 2 points)."""
    g = (1.0 + math.sqrt(5.0)) / 2.0
    v = [(-1, g, 0), (1, g, 0), (-1, -g, 0), (1, -g, 0), (0, -1, g), (0, 1, g), (0, -1, -g), (0, 1, -g),
         (g, 0, -1), (g, 0, 1), (-g, 0, -1), (-g, 0, 1)]
    f = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4), (11, 10, 2), (10, 7, 6),
         (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9), (4, 9, 5), (2, 4, 11), (6, 2, 10),
         (8, 6, 7), (9, 8, 1)]
    verts = np.asarray(v, dtype=np.float64)
    verts /= np.linalg.norm(verts, axis=1)[:, None]
    faces = np.asarray(f, dtype=np.int64)
    for _ in range(level):
        edges = np.concatenate([faces[:, [0, 1]], faces[:, [1, 2]], faces[:, [2, 0]]])
        edges.sort(axis=1)
        uniq, inv = np.unique(edges, axis=0, return_inverse=True)
        inv = inv.reshape(-1)
        mid = verts[uniq[:, 0]] + verts[uniq[:, 1]]
        mid /= np.linalg.norm(mid, axis=1)[:, None]
        base = verts.shape[0]
        verts = np.concatenate([verts, mid])
        nf = faces.shape[0]
        m01, m12, m20 = base + inv[:nf], base + inv[nf:2 * nf], base + inv[2 * nf:]
        a, b, c = faces[:, 0], faces[:, 1], faces[:, 2]
        faces = np.concatenate([np.stack([a, m01, m20], 1), np.stack([b, m12, m01], 1), np.stack([c, m20, m12], 1),
                                np.stack([m01, m12, m20], 1)])
    return verts


def jittered_voronoi(level: int, seed: int, jitter: float = 0.2):
    """Spherical Voronoi mesh of jittered level-`level` icosahedral generators.

    Each generator moves by N(0, (jitter * local spacing)^2) in its tangent
    plane (local spacing = mean distance to the six/five nearest vertices of
    the unjittered mesh, approximated by the level's mean edge length) and is
    reprojected to the sphere.  Cells are CCW seen from outside.
    """
    from scipy.spatial import SphericalVoronoi

    from . import fields as _fields

    pts = icosahedron_vertices(level)
    n = pts.shape[0]
    spacing = 1.1 * math.sqrt(4.0 * math.pi / n)
    z = _fields.normal(seed, 3 * n).reshape(n, 3) * (jitter * spacing)
    z -= (z * pts).sum(axis=1)[:, None] * pts
    pts = pts + z
    pts /= np.linalg.norm(pts, axis=1)[:, None]
    sv = SphericalVoronoi(pts, radius=1.0)
    sv.sort_vertices_of_regions()
    regions = sv.regions
    counts = np.fromiter((len(r) for r in regions), dtype=np.int64, count=len(regions))
    offsets = np.zeros(len(regions) + 1, dtype=np.int64)
    np.cumsum(counts, out=offsets[1:])
    indices = np.fromiter((v for r in regions for v in r), dtype=np.int64, count=int(offsets[-1]))
    vx = sv.vertices
    lon = np.degrees(np.arctan2(vx[:, 1], vx[:, 0]))
    lon = np.where(lon < 0.0, lon + 360.0, lon)
    lat = np.degrees(np.arctan2(vx[:, 2], np.hypot(vx[:, 0], vx[:, 1])))
    # orientation check per cell (sort_vertices_of_regions yields CCW; verify)
    p = _xyz_deg(lon, lat)
    cell_of = np.repeat(np.arange(len(regions)), counts)
    nxt = np.arange(indices.shape[0]) + 1
    nxt[offsets[1:] - 1] = offsets[:-1]
    cr = np.cross(p[indices], p[indices[nxt]])
    s = np.zeros(len(regions))
    np.add.at(s, cell_of, np.einsum("ij,ij->i", cr, pts[cell_of]))
    if np.any(s < 0):
        for c in np.nonzero(s < 0)[0]:
            b, e = offsets[c], offsets[c + 1]
            indices[b:e] = indices[b:e][::-1].copy()
    mesh = _mesh(lon, lat, offsets, indices)
    mesh["generators"] = pts
    return mesh


def cell_centroids_xyz(mesh):
    """Normalised mean of vertex xyz per cell (the cap-filter centroid rule)."""
    lon, lat = mesh["coords"][:, 0], mesh["coords"][:, 1]
    p = _xyz_deg(lon, lat)
    off, idx = mesh["offsets"], mesh["indices"]
    counts = np.diff(off)
    cell_of = np.repeat(np.arange(counts.shape[0]), counts)
    s = np.zeros((counts.shape[0], 3))
    np.add.at(s, cell_of, p[idx])
    s /= counts[:, None]
    return s / np.linalg.norm(s, axis=1)[:, None]


def latlon_cell_centres(mesh_ni: int, mesh_nj: int, res: float, min_x: float = 0.0, min_y: float = -90.0):
    """(lon, lat) degrees of regular lat-lon cell centres, cell order i + j*ni."""
    lon = min_x + (np.arange(mesh_ni) + 0.5) * res
    lat = min_y + (np.arange(mesh_nj) + 0.5) * res
    return np.tile(lon, mesh_nj), np.repeat(lat, mesh_ni)
