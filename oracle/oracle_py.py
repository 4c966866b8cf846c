"""ctypes binding of the CPU oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker.  The product library never
loads it.
"""
from __future__ import annotations

import ctypes as C
import os
import pathlib

import numpy as np

_HERE = pathlib.Path(__file__).resolve().parent
_LIB = None
_REF = None

_D = C.POINTER(C.c_double)
_F = C.POINTER(C.c_float)
_I64 = C.POINTER(C.c_int64)
_I32 = C.POINTER(C.c_int32)


def _p(a, t):
    return None if a is None else a.ctypes.data_as(t)


def lib():
    global _LIB
    if _LIB is None:
        path = _HERE / "liboracle.so"
        if not path.exists():
            raise RuntimeError(f"oracle not built: {path} (run make -C oracle)")
        L = C.CDLL(str(path))
        L.om_polygon_area.restype = C.c_double
        L.om_polygon_area.argtypes = [_D, C.c_int]
        L.om_overlap_area.restype = C.c_double
        L.om_overlap_area.argtypes = [_D, C.c_int, _D, C.c_int]
        L.om_clip.argtypes = [_D, C.c_int, _D, C.c_int, _D, C.POINTER(C.c_int)]
        L.om_lonlat_to_xyz.argtypes = [C.c_double, C.c_double, _D]
        L.om_detect_regular.argtypes = [_D, C.c_int64, _I64, _I64, C.c_int64, _D, _I64]
        L.om_detect_rectilinear.argtypes = [_D, C.c_int64, _I64, _I64, C.c_int64, _I64]
        L.om_cell_areas.argtypes = [_D, C.c_int64, _I64, _I64, C.c_int64, C.c_int, _D]
        L.om_degenerate.argtypes = [_D, C.c_int64, _I64, _I64, C.c_int64, _I32]
        L.om_cap_data.argtypes = [_D, C.c_int64, _I64, _I64, C.c_int64, C.c_int, _D, _D]
        L.om_cap_rejects.restype = C.c_int
        L.om_cap_rejects.argtypes = [_D, C.c_double, _D, C.c_double]
        L.om_aabbs.argtypes = [_D, C.c_int64, _I64, _I64, C.c_int64, C.c_int, _F]
        L.om_cell_centroids.argtypes = [_D, C.c_int64, _I64, _I64, C.c_int64, C.c_int, _F, _F]
        L.om_generate.restype = C.c_void_p
        L.om_generate.argtypes = [_D, C.c_int64, _I64, _I64, C.c_int64, _D, _I32, _D, C.c_int64, _I64, _I64, C.c_int64, _D,
                                  _I32, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int]
        L.om_result_error.restype = C.c_char_p
        L.om_result_error.argtypes = [C.c_void_p]
        L.om_result_nnz.restype = C.c_int64
        L.om_result_nnz.argtypes = [C.c_void_p]
        L.om_result_path.restype = C.c_int
        L.om_result_path.argtypes = [C.c_void_p]
        L.om_result_stats.argtypes = [C.c_void_p, _I64]
        L.om_result_ops.argtypes = [C.c_void_p, _D]
        L.om_result_csr.argtypes = [C.c_void_p, _I64, _I64, _D]
        L.om_result_fracs.argtypes = [C.c_void_p, _D, _D, _D, _D]
        L.om_result_n_candidates.restype = C.c_int64
        L.om_result_n_candidates.argtypes = [C.c_void_p]
        L.om_result_candidates.argtypes = [C.c_void_p, _I64, _I64]
        L.om_result_free.argtypes = [C.c_void_p]
        L.om_to_csr.argtypes = [_I64, _I64, _D, C.c_int64, C.c_int64, _I64, _I64, _D]
        L.om_batch_apply_csr.argtypes = [_I64, _I64, _D, C.c_int64, C.c_int64, _D, _D, C.c_int64, C.c_int]
        L.om_nan_batch_apply.argtypes = [_I64, _I64, _D, C.c_int64, C.c_int64, _D, _D, C.c_int64, C.c_double]
        L.om_masked_apply.argtypes = [_I64, _I64, _D, C.c_int64, C.c_int64, _D, _D, C.c_int64, _I32, C.c_int64]
        L.om_check_conservation.argtypes = [_D, C.c_int64, _D, C.c_int64, _D, _D, _D, _D, C.c_int, _D]
        L.om_max_threads.restype = C.c_int
        _LIB = L
    return _LIB


def ref_lib():
    """The reference's own spherical_geometry.hpp built under oracle/_ref (or None)."""
    global _REF
    if _REF is None:
        path = _HERE / "_ref" / "libref_geometry.so"
        if not path.exists():
            return None
        R = C.CDLL(str(path))
        R.ref_lonlat_to_xyz.argtypes = [C.c_double, C.c_double, _D]
        R.ref_polygon_area.restype = C.c_double
        R.ref_polygon_area.argtypes = [_D, C.c_int]
        R.ref_overlap_area.restype = C.c_double
        R.ref_overlap_area.argtypes = [_D, C.c_int, _D, C.c_int]
        _REF = R
    return _REF


def _xyz(a):
    return np.ascontiguousarray(a, dtype=np.float64).reshape(-1, 3)


def polygon_area(xyz) -> float:
    a = _xyz(xyz)
    return lib().om_polygon_area(_p(a, _D), a.shape[0])


def overlap_area(subj, clip) -> float:
    a, b = _xyz(subj), _xyz(clip)
    return lib().om_overlap_area(_p(a, _D), a.shape[0], _p(b, _D), b.shape[0])


def clip(subj, clip_poly) -> np.ndarray:
    a, b = _xyz(subj), _xyz(clip_poly)
    out = np.zeros((32, 3))
    n = C.c_int(0)
    lib().om_clip(_p(a, _D), a.shape[0], _p(b, _D), b.shape[0], _p(out, _D), C.byref(n))
    return out[: n.value].copy()


def lonlat_to_xyz(lon_rad: float, lat_rad: float) -> np.ndarray:
    out = np.zeros(3)
    lib().om_lonlat_to_xyz(lon_rad, lat_rad, _p(out, _D))
    return out


def _mesh_args(mesh):
    coords = np.ascontiguousarray(mesh["coords"].T).reshape(-1)  # layout_left flat: lons then lats
    off = np.ascontiguousarray(mesh["offsets"], dtype=np.int64)
    idx = np.ascontiguousarray(mesh["indices"], dtype=np.int64)
    return coords, off, idx


def detect_regular(mesh):
    coords, off, idx = _mesh_args(mesh)
    o6 = np.zeros(6)
    nij = np.zeros(2, dtype=np.int64)
    ok = lib().om_detect_regular(_p(coords, _D), mesh["coords"].shape[0], _p(off, _I64), _p(idx, _I64), off.shape[0] - 1,
                                 _p(o6, _D), _p(nij, _I64))
    return bool(ok), o6, nij


def detect_rectilinear(mesh):
    coords, off, idx = _mesh_args(mesh)
    nij = np.zeros(2, dtype=np.int64)
    ok = lib().om_detect_rectilinear(_p(coords, _D), mesh["coords"].shape[0], _p(off, _I64), _p(idx, _I64), off.shape[0] - 1,
                                     _p(nij, _I64))
    return bool(ok), nij


def cell_areas(mesh, coord_sys: int = 0) -> np.ndarray:
    coords, off, idx = _mesh_args(mesh)
    out = np.zeros(off.shape[0] - 1)
    lib().om_cell_areas(_p(coords, _D), mesh["coords"].shape[0], _p(off, _I64), _p(idx, _I64), off.shape[0] - 1, coord_sys,
                        _p(out, _D))
    return out


def degenerate(mesh) -> np.ndarray:
    coords, off, idx = _mesh_args(mesh)
    out = np.zeros(off.shape[0] - 1, dtype=np.int32)
    lib().om_degenerate(_p(coords, _D), mesh["coords"].shape[0], _p(off, _I64), _p(idx, _I64), off.shape[0] - 1, _p(out, _I32))
    return out


def cap_data(mesh, coord_sys: int = 0):
    coords, off, idx = _mesh_args(mesh)
    n = off.shape[0] - 1
    cen = np.zeros((n, 3))
    rad = np.zeros(n)
    lib().om_cap_data(_p(coords, _D), mesh["coords"].shape[0], _p(off, _I64), _p(idx, _I64), n, coord_sys, _p(cen, _D),
                      _p(rad, _D))
    return cen, rad


def cap_rejects(cs, rs, cd, rd) -> int:
    a = np.ascontiguousarray(cs, dtype=np.float64)
    b = np.ascontiguousarray(cd, dtype=np.float64)
    return lib().om_cap_rejects(_p(a, _D), float(rs), _p(b, _D), float(rd))


def aabbs(mesh, coord_sys: int = 0) -> np.ndarray:
    coords, off, idx = _mesh_args(mesh)
    n = off.shape[0] - 1
    out = np.zeros((n, 6), dtype=np.float32)
    lib().om_aabbs(_p(coords, _D), mesh["coords"].shape[0], _p(off, _I64), _p(idx, _I64), n, coord_sys, _p(out, _F))
    return out


class OracleError(RuntimeError):
    pass


def generate(src, dst, method: int = 2, norm_type: int = 0, unmapped_error: bool = False, brute: bool = False,
             keep_candidates: bool = False, coord_sys: int = 0, n_threads: int = 0, extrap: int = 1, parallel_clip: bool = False,
             count_ops: bool = False):
    """Weights as a dict (CSR + fractions + areas + stats); raises OracleError like the reference throws.

    parallel_clip: the all-parallel variant (every row's clips in parallel, then
    the sequential accumulation; same results).  count_ops: with the
    sequential loop, the FP64 operation count of the clips in out["ops"]."""
    sc, so, si = _mesh_args(src)
    dc, do, di = _mesh_args(dst)
    s_area = src.get("areas")
    d_area = dst.get("areas")
    s_mask = src.get("mask")
    d_mask = dst.get("mask")
    if s_mask is not None:
        s_mask = np.ascontiguousarray(s_mask, dtype=np.int32)
    if d_mask is not None:
        d_mask = np.ascontiguousarray(d_mask, dtype=np.int32)
    flags = (1 if brute else 0) | (2 if keep_candidates else 0) | (4 if parallel_clip else 0) | (8 if count_ops else 0)
    L = lib()
    h = L.om_generate(_p(sc, _D), src["coords"].shape[0], _p(so, _I64), _p(si, _I64), so.shape[0] - 1, _p(s_area, _D),
                      _p(s_mask, _I32), _p(dc, _D), dst["coords"].shape[0], _p(do, _I64), _p(di, _I64), do.shape[0] - 1,
                      _p(d_area, _D), _p(d_mask, _I32), coord_sys, method, norm_type, 1 if unmapped_error else 0, flags,
                      extrap, n_threads)
    try:
        err = L.om_result_error(h).decode()
        if err:
            raise OracleError(err)
        nnz = L.om_result_nnz(h)
        n_src, n_dst = so.shape[0] - 1, do.shape[0] - 1
        out = {
            "n_src": n_src, "n_dst": n_dst,
            "row_ptr": np.zeros(n_dst + 1, dtype=np.int64), "col": np.zeros(nnz, dtype=np.int64),
            "val": np.zeros(nnz), "frac_a": np.zeros(n_src), "frac_b": np.zeros(n_dst),
            "area_a": np.zeros(n_src), "area_b": np.zeros(n_dst),
        }
        L.om_result_csr(h, _p(out["row_ptr"], _I64), _p(out["col"], _I64), _p(out["val"], _D))
        L.om_result_fracs(h, _p(out["frac_a"], _D), _p(out["frac_b"], _D), _p(out["area_a"], _D), _p(out["area_b"], _D))
        st = np.zeros(3, dtype=np.int64)
        L.om_result_stats(h, _p(st, _I64))
        out["n_candidates"], out["n_cap_rejected"], out["n_clipped"] = (int(x) for x in st)
        out["path"] = ("general", "regular", "rectilinear")[L.om_result_path(h)]
        if count_ops:
            ops = np.zeros(5)
            L.om_result_ops(h, _p(ops, _D))
            out["ops"] = {"flops": float(ops[0]), "pairs": int(ops[1]), "edges": int(ops[2]), "intersections": int(ops[3]),
                          "fan_terms": int(ops[4])}
        if keep_candidates:
            nc = L.om_result_n_candidates(h)
            co = np.zeros(n_dst + 1, dtype=np.int64)
            ci = np.zeros(nc, dtype=np.int64)
            L.om_result_candidates(h, _p(co, _I64), _p(ci, _I64))
            out["cand_offsets"], out["cand_idx"] = co, ci
        return out
    finally:
        L.om_result_free(h)


def to_csr(rows, cols, vals, n_rows):
    rows = np.ascontiguousarray(rows, dtype=np.int64)
    cols = np.ascontiguousarray(cols, dtype=np.int64)
    vals = np.ascontiguousarray(vals, dtype=np.float64)
    rp = np.zeros(n_rows + 1, dtype=np.int64)
    c = np.zeros(rows.shape[0], dtype=np.int64)
    v = np.zeros(rows.shape[0])
    lib().om_to_csr(_p(rows, _I64), _p(cols, _I64), _p(vals, _D), rows.shape[0], n_rows, _p(rp, _I64), _p(c, _I64), _p(v, _D))
    return rp, c, v


def batch_apply(row_ptr, col, val, n_src, x, n_threads: int = 0):
    """FP64 Y = A X for layout_left X [n_src, k] (numpy array in Fortran order or 1-D)."""
    row_ptr = np.ascontiguousarray(row_ptr, dtype=np.int64)
    col = np.ascontiguousarray(col, dtype=np.int64)
    val = np.ascontiguousarray(val, dtype=np.float64)
    x = np.asarray(x, dtype=np.float64)
    k = 1 if x.ndim == 1 else x.shape[1]
    xf = np.asfortranarray(x.reshape(n_src, k))
    n_rows = row_ptr.shape[0] - 1
    y = np.zeros((n_rows, k), order="F")
    lib().om_batch_apply_csr(_p(row_ptr, _I64), _p(col, _I64), _p(val, _D), n_rows, n_src, xf.ctypes.data_as(_D),
                             y.ctypes.data_as(_D), k, n_threads)
    return y[:, 0].copy() if x.ndim == 1 else y


def nan_batch_apply(row_ptr, col, val, n_src, x, na_thres: float):
    """nan_batch_apply (apply.hpp:359-420) for layout_left X [n_src, k] or 1-D."""
    row_ptr = np.ascontiguousarray(row_ptr, dtype=np.int64)
    col = np.ascontiguousarray(col, dtype=np.int64)
    val = np.ascontiguousarray(val, dtype=np.float64)
    x = np.asarray(x, dtype=np.float64)
    k = 1 if x.ndim == 1 else x.shape[1]
    xf = np.asfortranarray(x.reshape(n_src, k))
    n_rows = row_ptr.shape[0] - 1
    y = np.zeros((n_rows, k), order="F")
    lib().om_nan_batch_apply(_p(row_ptr, _I64), _p(col, _I64), _p(val, _D), n_rows, n_src, xf.ctypes.data_as(_D),
                             y.ctypes.data_as(_D), k, C.c_double(na_thres))
    return y[:, 0].copy() if x.ndim == 1 else y


def masked_apply(row_ptr, col, val, n_src, x, y0, mask):
    """Masked apply (apply.hpp:438-484): rows with mask 0 keep y0's values."""
    row_ptr = np.ascontiguousarray(row_ptr, dtype=np.int64)
    col = np.ascontiguousarray(col, dtype=np.int64)
    val = np.ascontiguousarray(val, dtype=np.float64)
    x = np.asarray(x, dtype=np.float64)
    k = 1 if x.ndim == 1 else x.shape[1]
    xf = np.asfortranarray(x.reshape(n_src, k))
    n_rows = row_ptr.shape[0] - 1
    y = np.asfortranarray(np.array(y0, dtype=np.float64).reshape(n_rows, k))
    m = None if mask is None else np.ascontiguousarray(mask, dtype=np.int32)
    lib().om_masked_apply(_p(row_ptr, _I64), _p(col, _I64), _p(val, _D), n_rows, n_src, xf.ctypes.data_as(_D), y.ctypes.data_as(_D), k,
                          None if m is None else _p(m, _I32), 0 if m is None else m.shape[0])
    return y[:, 0].copy() if x.ndim == 1 else y


def check_conservation(src_field, dst_field, w, norm_type: int = 0):
    out = np.zeros(4)
    s = np.ascontiguousarray(src_field, dtype=np.float64)
    d = np.ascontiguousarray(dst_field, dtype=np.float64)
    lib().om_check_conservation(_p(s, _D), s.shape[0], _p(d, _D), d.shape[0], _p(w["area_a"], _D), _p(w["frac_a"], _D),
                                _p(w["area_b"], _D), _p(w["frac_b"], _D), norm_type, _p(out, _D))
    return {"src_integral": out[0], "dst_integral": out[1], "absolute_error": out[2], "relative_error": out[3]}


def max_threads() -> int:
    return lib().om_max_threads()


def available() -> bool:
    return (_HERE / "liboracle.so").exists()


os.environ.setdefault("OMP_PROC_BIND", "false")


def cell_centroids(mesh, coord_sys: int = 0):
    """compute_cell_centroids_xy rounded to float: (lon, lat) float32 arrays."""
    c, o, i = _mesh_args(mesh)
    n = o.shape[0] - 1
    lon = np.zeros(n, dtype=np.float32)
    lat = np.zeros(n, dtype=np.float32)
    lib().om_cell_centroids(_p(c, _D), mesh["coords"].shape[0], _p(o, _I64), _p(i, _I64), n, coord_sys, _p(lon, _F), _p(lat, _F))
    return lon, lat


def halo_pattern(cols, src_global_ids, owner_of_src):
    """HaloPattern of one rank (weight_generator.cpp:4525-4680), restated over
    the local matrix's column indices: the rank is the owner of the first
    local source cell; used local columns whose global id another rank owns
    are grouped by owner rank, ascending local column within a rank (the
    reference takes first appearance in its COO order, then an unstable sort
    by rank); slot g is remapped to column n_local + g.  Returns
    (pattern dict, remapped columns)."""
    cols = np.asarray(cols, dtype=np.int64)
    gid = np.asarray(src_global_ids, dtype=np.int64)
    owner = np.asarray(owner_of_src, dtype=np.int32)
    n_local = gid.shape[0]
    me = int(owner[gid[0]]) if n_local > 0 and 0 <= gid[0] < owner.shape[0] else -1
    used = np.unique(cols[(cols >= 0) & (cols < n_local)])
    g = gid[used]
    ok = (g >= 0) & (g < owner.shape[0])
    used, g = used[ok], g[ok]
    off = used[owner[g] != me]
    ranks = owner[gid[off]]
    order = np.argsort(ranks, kind="stable")
    off, ranks = off[order], ranks[order]
    src_ranks, starts = np.unique(ranks, return_index=True)
    pattern = {"source_ranks": src_ranks.astype(np.int32),
               "rank_offsets": np.concatenate([starts, [off.shape[0]]]).astype(np.int64),
               "needed_global_src_ids": gid[off], "gather_slot": np.arange(off.shape[0], dtype=np.int64)}
    remap = np.arange(n_local, dtype=np.int64)
    remap[off] = n_local + np.arange(off.shape[0])
    return pattern, remap[cols]


# ----------------------------------------------------------------- mesh ingest
def rule_regular_latlon(min_x, max_x, min_y, max_y, r_x, r_y):
    """RuleGenerator RegularLatLon (rule_generator.cpp:105-169): ni =
    floor((max_x - min_x) / r_x), nj likewise; node idx = i + j (ni + 1) at
    (min_x + i r_x, min_y + j r_y); cell c = i + j ni with corners BL, BR, TR,
    TL.  Returns (coords [n_nodes, 2], offsets, indices)."""
    ni = int(np.floor((max_x - min_x) / r_x))
    nj = int(np.floor((max_y - min_y) / r_y))
    nodes = []
    for idx in range((ni + 1) * (nj + 1)):
        i, j = idx % (ni + 1), idx // (ni + 1)
        nodes.append((min_x + float(i) * r_x, min_y + float(j) * r_y))
    ind = []
    for c in range(ni * nj):
        i, j, st = c % ni, c // ni, ni + 1
        ind += [j * st + i, j * st + i + 1, (j + 1) * st + i + 1, (j + 1) * st + i]
    return np.array(nodes, dtype=np.float64).reshape(-1, 2), np.arange(0, 4 * ni * nj + 1, 4, dtype=np.int64), np.array(ind, dtype=np.int64)


def structured_to_unstructured(ni, nj, clon, clat, corner_lon=None, corner_lat=None):
    """StructuredGrid(ni, nj, centres).to_unstructured()
    (structured_grid.cpp:23-83 constructor, :107-174 corner synthesis,
    :209-272 conversion): with no corners each corner (ci, cj) is the mean of
    the centres of the existing cells (ci + di, cj + dj), dj then di over
    {-1, 0}; a row spanning 350-370 degrees (ni |clon[1] - clon[0]|) wraps, the
    wrapped cell's longitude moved by -360 (left) or +360 (right).  Node
    ci + cj (ni + 1); cell i + j ni with corners BL, BR, TR, TL."""
    clon = np.asarray(clon, dtype=np.float64)
    clat = np.asarray(clat, dtype=np.float64)
    if corner_lon is None:
        periodic = ni > 1 and 350.0 < ni * abs(clon[1] - clon[0]) < 370.0
        corner_lon = np.zeros((ni + 1) * (nj + 1))
        corner_lat = np.zeros((ni + 1) * (nj + 1))
        for cj in range(nj + 1):
            for ci in range(ni + 1):
                sx = sy = 0.0
                n = 0
                for dj in (-1, 0):
                    for di in (-1, 0):
                        i, j, off = ci + di, cj + dj, 0.0
                        if periodic:
                            if i < 0:
                                i, off = ni - 1, -360.0
                            elif i >= ni:
                                i, off = 0, 360.0
                        elif not 0 <= i < ni:
                            continue
                        if not 0 <= j < nj:
                            continue
                        sx += clon[i + j * ni] + off
                        sy += clat[i + j * ni]
                        n += 1
                corner_lon[ci + cj * (ni + 1)] = sx / n
                corner_lat[ci + cj * (ni + 1)] = sy / n
    coords = np.stack([np.asarray(corner_lon, dtype=np.float64), np.asarray(corner_lat, dtype=np.float64)], 1)
    ind = []
    for c in range(ni * nj):
        i, j, st = c % ni, c // ni, ni + 1
        ind += [i + j * st, i + 1 + j * st, i + 1 + (j + 1) * st, i + (j + 1) * st]
    return coords, np.arange(0, 4 * ni * nj + 1, 4, dtype=np.int64), np.array(ind, dtype=np.int64)
