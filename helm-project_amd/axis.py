"""ctypes binding of libaxis_hip.so (the C ABI declared in include/axis_hip.h).

Mirrors the reference's interop layer (libs/axis/src/fortran/axis_c_interop.cpp)
and solver API: meshes and matrices are integer handles, every call that
returns AXIS_ERROR raises AxisError carrying the library's message.

Fields are column-major [n, k] (layout_left): in numpy a Fortran-ordered array
or a 1-D vector; in torch a row-major tensor of shape (k, n), so each field is
one contiguous row.  Device tensors are passed by pointer (no copy).
"""
from __future__ import annotations

import ctypes as C
import os
import pathlib

import numpy as np

_HERE = pathlib.Path(__file__).resolve().parent
# AXIS_HIP_LIBRARY selects another build of the same library (e.g. a debug build)
LIB_PATH = pathlib.Path(os.environ.get("AXIS_HIP_LIBRARY", _HERE / "lib" / "libaxis_hip.so"))

F64, F32, BF16 = 0, 1, 2
PLAIN, MIXED, MIXED_ACC64 = 0, 1, 2
BILINEAR, NEAREST, CONSERVATIVE = 0, 1, 2
# InterpolationMethod declaration order (regrid_config.hpp:24-33)
M_BILINEAR, M_NEAREST, M_BICUBIC, M_PATCH, M_CONSERVATIVE1, M_CONSERVATIVE2 = range(6)

_D = C.POINTER(C.c_double)
_I64 = C.POINTER(C.c_int64)
_I32 = C.POINTER(C.c_int32)
_lib = None


class AxisError(RuntimeError):
    pass


# int (*)(const int64_t *needed_global_src_ids, int64_t n_remote, double *out, void *user)
GATHER_FN = C.CFUNCTYPE(C.c_int, C.POINTER(C.c_int64), C.c_int64, C.POINTER(C.c_double), C.c_void_p)


class RegridConfig(C.Structure):
    """axis_regrid_config; defaults are RegridConfig's (regrid_config.hpp:80-89)."""

    _fields_ = [("method", C.c_int), ("norm_type", C.c_int), ("line_type", C.c_int), ("unmapped", C.c_int),
                ("use_limiter", C.c_int), ("extrap_method", C.c_int)]

    def __init__(self, method=M_BILINEAR, norm_type=0, line_type=1, unmapped=1, use_limiter=0, extrap_method=1):
        super().__init__(method, norm_type, line_type, unmapped, use_limiter, extrap_method)


class _StdString(C.Structure):
    """A libstdc++ std::string as the C++ side reads it: data pointer, length,
    then a 16-byte buffer that holds short strings (the pointer then points at
    it) or the capacity of a longer one."""

    _fields_ = [("p", C.c_void_p), ("n", C.c_size_t), ("buf", C.c_char * 16)]

    def set(self, text: str, keep: list):
        b = text.encode()
        self.n = len(b)
        if len(b) < 16:
            self.buf = b
            self.p = C.addressof(self) + _StdString.buf.offset
        else:
            mem = C.create_string_buffer(b)
            keep.append(mem)
            self.p = C.addressof(mem)
            C.cast(C.addressof(self) + _StdString.buf.offset, C.POINTER(C.c_size_t))[0] = len(b)


class _View1(C.Structure):  # field_view<T, 1>: data handle, extent
    _fields_ = [("data", C.c_void_p), ("n", C.c_size_t)]


class _View2(C.Structure):  # field_view<T, 2> (layout_left): data handle, extents
    _fields_ = [("data", C.c_void_p), ("n0", C.c_size_t), ("n1", C.c_size_t)]


class GridDescriptor(C.Structure):
    """axis::ingest::GridDescriptor (ingest/grid_descriptor.hpp:37-216) laid
    out field by field in its declaration order, for
    axis_mesh_from_descriptor_c.  Build with grid_descriptor()."""

    _fields_ = [
        ("kind", C.c_uint8), ("coord_system", C.c_uint8),
        # CfParams
        ("cf_grid_mapping_name", _StdString), ("cf_earth_radius", C.c_double), ("cf_semi_major", C.c_double),
        ("cf_semi_minor", C.c_double), ("cf_inverse_flattening", C.c_double),
        # UgridParams
        ("ugrid_topology_dimension", C.c_int), ("ugrid_start_index", C.c_int), ("ugrid_mesh_name", _StdString),
        # GribParams
        ("grib_grid_type", _StdString), ("grib_ni", C.c_int64), ("grib_nj", C.c_int64), ("grib_gaussian_n", C.c_int64),
        ("grib_scanning_mode", C.c_uint8), ("grib_earth_shape_code", C.c_int),
        # ProjectedParams, NamedGridParams
        ("proj_string", _StdString), ("named_grid_name", _StdString),
        # GridRulesParams
        ("rules_kind", _StdString), ("rules_min_x", C.c_double), ("rules_max_x", C.c_double), ("rules_min_y", C.c_double),
        ("rules_max_y", C.c_double), ("rules_r_x", C.c_double), ("rules_r_y", C.c_double), ("rules_gaussian_n", C.c_int64),
        ("rules_proj_string", _StdString),
        # BufferViews
        ("center_x", _View1), ("center_y", _View1), ("corner_x", _View1), ("corner_y", _View1), ("node_coords", _View2),
        ("conn_offsets", _View1), ("conn_indices", _View1), ("cell_areas", _View1), ("cell_mask", _View1),
        ("ni", C.c_size_t), ("nj", C.c_size_t),
    ]


# ConventionKind (grid_descriptor.hpp:37-44)
DESC_CF, DESC_UGRID, DESC_GRIB, DESC_PROJECTED, DESC_NAMED, DESC_RULES = range(6)


def grid_descriptor(kind: int, coord_system: int = 0, strings=None, scalars=None, arrays=None):
    """A GridDescriptor with the reference's defaults, then the given string
    fields (e.g. {"rules_kind": "RegularLatLon"}), scalar fields and buffers
    (numpy arrays: 1-D for the views, [n, ndim] for node_coords).  The arrays
    are kept alive on the returned object."""
    d = GridDescriptor()
    keep = []
    d.kind, d.coord_system = kind, coord_system
    d.ugrid_topology_dimension = 2
    d.rules_min_x, d.rules_max_x, d.rules_min_y, d.rules_max_y, d.rules_r_x, d.rules_r_y = -180.0, 180.0, -90.0, 90.0, 1.0, 1.0
    for name, _ in GridDescriptor._fields_:
        if name not in ("kind", "coord_system") and isinstance(getattr(d, name), _StdString):
            getattr(d, name).set("", keep)
    for name, text in (strings or {}).items():
        getattr(d, name).set(text, keep)
    for name, v in (scalars or {}).items():
        setattr(d, name, v)
    for name, a in (arrays or {}).items():
        if name == "node_coords":
            a = np.asfortranarray(a, dtype=np.float64)
            d.node_coords.data, d.node_coords.n0, d.node_coords.n1 = a.ctypes.data, a.shape[0], a.shape[1]
        else:
            dt = {"conn_offsets": np.int64, "conn_indices": np.int64, "cell_mask": np.int32}.get(name, np.float64)
            a = np.ascontiguousarray(a, dtype=dt)
            v = getattr(d, name)
            v.data, v.n = a.ctypes.data, a.shape[0]
        keep.append(a)
    d._keep = keep
    return d


def lib():
    """Loads the library; raises if it was not built (there is no fallback)."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise AxisError(f"{LIB_PATH} is missing: build it with `make -C helm-project_amd` (no CPU fallback exists)")
    L = C.CDLL(str(LIB_PATH))
    sig = {
        "axis_init_c": [],
        "axis_mesh_create_c": [_D, C.c_int64, C.c_int64, _I64, C.c_int64, _I64, C.c_int, _D, _I32, C.POINTER(C.c_int)],
        "axis_mesh_regular_latlon_c": [C.c_double] * 6 + [C.POINTER(C.c_int)],
        "axis_mesh_from_descriptor_c": [C.c_void_p, C.POINTER(C.c_int)],
        "axis_mesh_from_named_c": [C.c_char_p, C.POINTER(C.c_int)],
        "axis_mesh_arrays_c": [C.c_int, _D, _I64, _I64],
        "axis_generate_vector_weights_c": [C.c_int, C.c_int, _D, _D, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int)],
        "axis_mesh_cell_centroids_c": [C.c_int, C.POINTER(C.c_float), C.POINTER(C.c_float), _I64],
        "axis_mesh_info_c": [C.c_int, _I64, _I64],
        "axis_mesh_grid_info_c": [C.c_int, C.POINTER(C.c_int), _I64, _D],
        "axis_generate_weights_c": [C.c_int, C.c_int, C.c_int, C.POINTER(C.c_int)],
        "axis_generate_weights_config_c": [C.c_int, C.c_int, C.POINTER(RegridConfig), C.POINTER(C.c_int)],
        "axis_matrix_from_coo_c": [_I64, _I64, _D, C.c_int64, C.c_int64, C.c_int64, _D, _D, _D, _D, C.POINTER(C.c_int)],
        "axis_matrix_to_csr_c": [C.c_int],
        "axis_matrix_serialize_c": [C.c_int, C.c_void_p, C.c_uint64, C.POINTER(C.c_uint64)],
        "axis_nan_batch_apply_c": [C.c_int, C.c_void_p, C.c_void_p, C.c_int64, C.c_int64, C.c_int64, C.c_double, C.c_void_p],
        "axis_apply_masked_c": [C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64, C.c_int64, C.c_int64, C.c_int64, C.c_void_p],
        "axis_generate_weights_distributed_c": [C.c_int, C.c_int, C.c_void_p, _I64, C.c_int64, _I64, C.c_int64, _I32, C.c_int64,
                                                C.POINTER(C.c_int), C.POINTER(C.c_int)],
        "axis_halo_pattern_info_c": [C.c_int, _I64, _I64],
        "axis_halo_pattern_arrays_c": [C.c_int, _I32, _I64, _I64, _I64],
        "axis_destroy_halo_pattern_c": [C.c_int],
        "axis_apply_distributed_c": [C.c_int, C.c_int, C.c_void_p, C.c_int64, C.c_void_p, C.c_int64, C.c_void_p, C.c_int64, C.c_int64,
                                     C.c_void_p],
        "axis_apply_distributed_callback_c": [C.c_int, C.c_int, C.c_void_p, C.c_int64, GATHER_FN, C.c_void_p, C.c_void_p, C.c_int64,
                                              C.c_void_p],
        "axis_matrix_deserialize_c": [C.c_void_p, C.c_uint64, C.POINTER(C.c_int)],
        "axis_matrix_write_esmf_c": [C.c_int, C.c_char_p],
        "axis_matrix_read_esmf_c": [C.c_char_p, C.POINTER(C.c_int)],
        "axis_matrix_info_c": [C.c_int, _I64, _I64, _I64],
        "axis_matrix_csr_c": [C.c_int, C.c_void_p, C.c_void_p, C.c_void_p],
        "axis_matrix_fractions_c": [C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p],
        "axis_matrix_stats_c": [C.c_int, _I64, _D],
        "axis_matrix_phase_ms_c": [C.c_int, _D, C.c_int],
        "axis_apply_c": [C.c_int, C.c_void_p, C.c_void_p, C.c_int, C.c_int],
        "axis_batch_apply_c": [C.c_int, C.c_void_p, C.c_int, C.c_void_p, C.c_int, C.c_int64, C.c_int64, C.c_int64, C.c_int,
                               C.c_void_p],
        "axis_apply_set_shape_c": [C.c_int, C.c_int64, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int],
        "axis_batch_apply_noalloc_c": [C.c_int, C.c_void_p, C.c_int, C.c_void_p, C.c_int, C.c_int64, C.c_int64, C.c_int64,
                                       C.c_int, C.c_void_p],
        "axis_batch_apply_async_c": [C.c_int, C.c_void_p, C.c_int, C.c_void_p, C.c_int, C.c_int64, C.c_int64, C.c_int64,
                                     C.c_int, C.c_void_p],
        "axis_autotune_apply_c": [C.c_int, C.c_void_p, C.c_int, C.c_void_p, C.c_int, C.c_int64, C.c_int, C.c_void_p,
                                  C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int)],
        "axis_check_conservation_c": [C.c_int, C.c_void_p, C.c_void_p, C.c_int, _D],
        "axis_destroy_mesh_c": [C.c_int],
        "axis_destroy_matrix_c": [C.c_int],
        "axis_batch_plan_create_c": [C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_void_p), C.POINTER(C.c_void_p),
                                     C.POINTER(C.c_int64), C.c_int, C.c_int, C.c_int, C.c_int, C.c_void_p,
                                     C.POINTER(C.c_int)],
        "axis_batch_plan_run_c": [C.c_int, C.c_void_p, C.c_int],
        "axis_batch_plan_info_c": [C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int), _I64, C.POINTER(C.c_int)],
        "axis_destroy_batch_plan_c": [C.c_int],
        "axis_mlp_create_c": [C.POINTER(C.c_float), C.c_uint64, C.c_float, C.POINTER(C.c_int)],
        "axis_mlp_params_c": [C.c_int, C.POINTER(C.c_float), C.POINTER(C.c_float)],
        "axis_mlp_set_params_c": [C.c_int, C.POINTER(C.c_float)],
        "axis_mlp_forward_c": [C.c_int, C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p],
        "axis_mlp_gradient_c": [C.c_int, C.c_void_p, C.c_void_p, C.c_int64, C.c_float, C.c_float, C.c_void_p,
                                C.POINTER(C.c_float), C.c_void_p],
        "axis_mlp_train_c": [C.c_int, C.c_void_p, C.c_void_p, C.c_int64, C.c_int64, C.c_float, C.c_float, C.c_float,
                             C.POINTER(C.c_float), C.c_void_p],
        "axis_destroy_mlp_c": [C.c_int],
        "axis_generate_weights_proposed_c": [C.c_int, C.c_int, C.POINTER(RegridConfig), C.c_int, C.c_double,
                                             C.POINTER(C.c_int), _I64],
        "axis_proposer_training_data_c": [C.c_int, C.c_int, C.c_float, C.c_void_p, C.c_void_p, _I64],
    }
    for name, args in sig.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = C.c_int
    L.axis_last_error_c.argtypes = []
    L.axis_last_error_c.restype = C.c_char_p
    _lib = L
    return L


def exported_symbols():
    """Names of the C entry points the binding declares."""
    return ["axis_init_c", "axis_last_error_c", "axis_mesh_create_c", "axis_mesh_regular_latlon_c", "axis_mesh_info_c", "axis_mesh_grid_info_c",
            "axis_mesh_cell_centroids_c",
            "axis_generate_weights_c", "axis_generate_weights_config_c", "axis_matrix_from_coo_c",
            "axis_matrix_to_csr_c", "axis_matrix_info_c", "axis_matrix_csr_c", "axis_matrix_fractions_c",
            "axis_matrix_stats_c", "axis_matrix_phase_ms_c", "axis_apply_c", "axis_batch_apply_c", "axis_batch_apply_async_c",
            "axis_autotune_apply_c", "axis_check_conservation_c", "axis_destroy_mesh_c", "axis_destroy_matrix_c",
            "axis_batch_plan_create_c", "axis_batch_plan_run_c", "axis_batch_plan_info_c", "axis_destroy_batch_plan_c",
            "axis_mlp_create_c", "axis_mlp_params_c", "axis_mlp_set_params_c", "axis_mlp_forward_c",
            "axis_mlp_gradient_c", "axis_mlp_train_c", "axis_destroy_mlp_c", "axis_generate_weights_proposed_c",
            "axis_proposer_training_data_c", "axis_matrix_serialize_c", "axis_matrix_deserialize_c",
            "axis_matrix_write_esmf_c", "axis_matrix_read_esmf_c", "axis_nan_batch_apply_c", "axis_apply_masked_c",
            "axis_generate_weights_distributed_c", "axis_halo_pattern_info_c", "axis_halo_pattern_arrays_c",
            "axis_destroy_halo_pattern_c", "axis_apply_distributed_c", "axis_apply_distributed_callback_c",
            "axis_mesh_from_descriptor_c", "axis_mesh_from_named_c", "axis_mesh_arrays_c", "axis_generate_vector_weights_c",
            "axis_batch_apply_noalloc_c", "axis_apply_set_shape_c"]


def _check(rc: int):
    if rc != 0:
        raise AxisError(lib().axis_last_error_c().decode())


def _ptr(a):
    """Data pointer of a numpy array or torch tensor (None -> NULL)."""
    if a is None:
        return None
    if isinstance(a, np.ndarray):
        return a.ctypes.data
    return a.data_ptr()


def _stream(stream):
    if stream is None:
        try:
            import torch

            if torch.cuda.is_available():
                return C.c_void_p(torch.cuda.current_stream().cuda_stream)
        except ImportError:
            pass
        return None
    return C.c_void_p(stream)


def init():
    _check(lib().axis_init_c())


class Mesh:
    """A mesh handle (UnstructuredMesh in HBM)."""

    def __init__(self, mesh: dict, coord_sys: int = 0):
        coords = np.ascontiguousarray(np.asarray(mesh["coords"]).T).reshape(-1)  # lon block, lat block
        off = np.ascontiguousarray(mesh["offsets"], dtype=np.int64)
        idx = np.ascontiguousarray(mesh["indices"], dtype=np.int64)
        areas = mesh.get("areas")
        mask = mesh.get("mask")
        areas = None if areas is None else np.ascontiguousarray(areas, dtype=np.float64)
        mask = None if mask is None else np.ascontiguousarray(mask, dtype=np.int32)
        h = C.c_int(0)
        _check(lib().axis_mesh_create_c(coords.ctypes.data_as(_D), coords.shape[0] // 2, 2, off.ctypes.data_as(_I64),
                                        off.shape[0] - 1, idx.ctypes.data_as(_I64), coord_sys,
                                        None if areas is None else areas.ctypes.data_as(_D),
                                        None if mask is None else mask.ctypes.data_as(_I32), C.byref(h)))
        self.handle = h.value

    @classmethod
    def regular_latlon(cls, min_x, max_x, min_y, max_y, r_x, r_y):
        self = cls.__new__(cls)
        h = C.c_int(0)
        _check(lib().axis_mesh_regular_latlon_c(min_x, max_x, min_y, max_y, r_x, r_y, C.byref(h)))
        self.handle = h.value
        return self

    @classmethod
    def from_descriptor(cls, desc: "GridDescriptor"):
        """MeshFactory::from_descriptor through axis_mesh_from_descriptor_c."""
        self = cls.__new__(cls)
        h = C.c_int(0)
        _check(lib().axis_mesh_from_descriptor_c(C.addressof(desc), C.byref(h)))
        self.handle = h.value
        return self

    @classmethod
    def from_named(cls, name: str):
        self = cls.__new__(cls)
        h = C.c_int(0)
        _check(lib().axis_mesh_from_named_c(name.encode(), C.byref(h)))
        self.handle = h.value
        return self

    def info(self):
        nn, nc = C.c_int64(0), C.c_int64(0)
        _check(lib().axis_mesh_info_c(self.handle, C.byref(nn), C.byref(nc)))
        return nn.value, nc.value

    def arrays(self):
        """(coords [n_nodes, 2], offsets, indices) copied out of the mesh."""
        nn, nc = self.info()
        coords = np.zeros(2 * nn)
        off = np.zeros(nc + 1, dtype=np.int64)
        _check(lib().axis_mesh_arrays_c(self.handle, coords.ctypes.data_as(_D), off.ctypes.data_as(_I64), None))
        idx = np.zeros(max(int(off[-1]), 1), dtype=np.int64)
        _check(lib().axis_mesh_arrays_c(self.handle, None, None, idx.ctypes.data_as(_I64)))
        return coords.reshape(2, nn).T, off, idx[:int(off[-1])]

    def grid_info(self):
        """(kind, (ni, nj), out6): kind 0 unstructured, 1 rectilinear, 2 regular;
        out6 = (lon_min, lon_max, delta_lon, lat_min, lat_max, delta_lat) when regular."""
        kind = C.c_int(0)
        ninj = np.zeros(2, dtype=np.int64)
        out6 = np.zeros(6)
        _check(lib().axis_mesh_grid_info_c(self.handle, C.byref(kind), ninj.ctypes.data_as(_I64), out6.ctypes.data_as(_D)))
        return kind.value, ninj, out6

    def cell_centroids(self):
        """(lon, lat, n_host): float32 cell centroids as the masked-source
        extrapolation uses them, and how many cells were recomputed on the host."""
        _, nc = self.info()
        lon = np.zeros(nc, dtype=np.float32)
        lat = np.zeros(nc, dtype=np.float32)
        nh = C.c_int64(0)
        _check(lib().axis_mesh_cell_centroids_c(self.handle, lon.ctypes.data_as(C.POINTER(C.c_float)),
                                                lat.ctypes.data_as(C.POINTER(C.c_float)), C.byref(nh)))
        return lon, lat, nh.value

    def close(self):
        if getattr(self, "handle", 0):
            _check(lib().axis_destroy_mesh_c(self.handle))
            self.handle = 0

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Matrix:
    """An InterpolationMatrix handle (CSR in HBM)."""

    def __init__(self, handle: int):
        self.handle = handle

    def info(self):
        a, b, c = C.c_int64(0), C.c_int64(0), C.c_int64(0)
        _check(lib().axis_matrix_info_c(self.handle, C.byref(a), C.byref(b), C.byref(c)))
        return a.value, b.value, c.value

    @property
    def n_src(self):
        return self.info()[0]

    @property
    def n_dst(self):
        return self.info()[1]

    @property
    def nnz(self):
        return self.info()[2]

    def to_csr(self):
        _check(lib().axis_matrix_to_csr_c(self.handle))

    def to_bytes(self) -> bytes:
        """The AXWC weight blob (WeightCache::serialize)."""
        n = C.c_uint64(0)
        _check(lib().axis_matrix_serialize_c(self.handle, None, 0, C.byref(n)))
        buf = C.create_string_buffer(n.value)
        _check(lib().axis_matrix_serialize_c(self.handle, buf, n.value, C.byref(n)))
        return buf.raw[:n.value]

    @classmethod
    def from_bytes(cls, data: bytes) -> "Matrix":
        """A matrix from an AXWC weight blob (WeightCache::deserialize)."""
        h = C.c_int(0)
        _check(lib().axis_matrix_deserialize_c(data, len(data), C.byref(h)))
        return cls(h.value)

    def csr(self):
        n_src, n_dst, nnz = self.info()
        rp = np.zeros(n_dst + 1, dtype=np.int64)
        col = np.zeros(nnz, dtype=np.int64)
        val = np.zeros(nnz)
        _check(lib().axis_matrix_csr_c(self.handle, rp.ctypes.data, col.ctypes.data, val.ctypes.data))
        return rp, col, val

    def fractions(self):
        n_src, n_dst, _ = self.info()
        fa, fb, aa, ab = np.zeros(n_src), np.zeros(n_dst), np.zeros(n_src), np.zeros(n_dst)
        _check(lib().axis_matrix_fractions_c(self.handle, fa.ctypes.data, fb.ctypes.data, aa.ctypes.data, ab.ctypes.data))
        return {"frac_a": fa, "frac_b": fb, "area_a": aa, "area_b": ab}

    def stats(self):
        out = np.zeros(4, dtype=np.int64)
        ms = C.c_double(0)
        _check(lib().axis_matrix_stats_c(self.handle, out.ctypes.data_as(_I64), C.byref(ms)))
        return {"candidates": int(out[0]), "host_arbitrated": int(out[1]), "cap_ties": int(out[2]),
                "max_row_nnz": int(out[3]), "gen_ms": ms.value}

    def phase_ms(self):
        """GPU time of the generation's phases (axis_matrix_phase_ms_c); None
        for a phase that did not run."""
        out = np.zeros(4, dtype=np.float64)
        _check(lib().axis_matrix_phase_ms_c(self.handle, out.ctypes.data_as(_D), 4))
        names = ("broad_phase_query", "cap_filter_kernel", "clip_kernel", "csr_assembly")
        return {k: (None if v < 0 else float(v)) for k, v in zip(names, out)}

    def apply(self, src, dst=None):
        """FP64 dst = S src through axis_apply_c (host numpy or device torch)."""
        n_src, n_dst, _ = self.info()
        if dst is None:
            if isinstance(src, np.ndarray):
                dst = np.zeros(n_dst)
            else:
                import torch

                dst = torch.empty(n_dst, dtype=torch.float64, device=src.device)
        if isinstance(src, np.ndarray):
            src = np.ascontiguousarray(src, dtype=np.float64)
        _check(lib().axis_apply_c(self.handle, _ptr(src), _ptr(dst), int(src.shape[0]), int(dst.shape[0])))
        return dst

    def batch_apply(self, src, dst, n_fields, src_dtype, dst_dtype, mode, stream=None, sync=True):
        """Device fields (torch (k, n) tensors or raw pointers)."""
        n_src, n_dst, _ = self.info()
        fn = lib().axis_batch_apply_c if sync else lib().axis_batch_apply_async_c
        _check(fn(self.handle, _ptr(src) if not isinstance(src, int) else src, src_dtype,
                  _ptr(dst) if not isinstance(dst, int) else dst, dst_dtype, n_src, n_dst, n_fields, mode,
                  _stream(stream)))
        return dst

    def nan_batch_apply(self, src, dst, n_fields, na_thres, stream=None):
        """nan_batch_apply on FP64 device fields ((k, n) torch tensors): NaN sources
        skipped, rows renormalised by their valid weight."""
        n_src, n_dst, _ = self.info()
        _check(lib().axis_nan_batch_apply_c(self.handle, _ptr(src), _ptr(dst), n_src, n_dst, n_fields, na_thres, _stream(stream)))
        return dst

    def apply_masked(self, src, dst, dst_mask, n_fields=1, stream=None):
        """Apply to FP64 device fields leaving rows with dst_mask == 0 unchanged
        (the mask applies when it has n_dst entries)."""
        n_src, n_dst, _ = self.info()
        n_mask = 0 if dst_mask is None else int(dst_mask.numel() if hasattr(dst_mask, "numel") else len(dst_mask))
        _check(lib().axis_apply_masked_c(self.handle, _ptr(src), _ptr(dst), _ptr(dst_mask), n_mask, n_src, n_dst, n_fields,
                                         _stream(stream)))
        return dst

    def apply_distributed(self, pattern, local_src, halo, dst, n_fields=1, stream=None):
        """Distributed apply with the gathered halo buffer: dst = S [local_src ; halo]
        (FP64 device fields, layout_left [n, n_fields])."""
        _, n_dst, _ = self.info()
        n_local = (local_src.numel() // n_fields) if local_src is not None else 0
        n_halo = (halo.numel() // n_fields) if halo is not None else 0
        _check(lib().axis_apply_distributed_c(self.handle, pattern.handle, _ptr(local_src), n_local, _ptr(halo), n_halo, _ptr(dst),
                                              dst.numel() // n_fields, n_fields, _stream(stream)))
        return dst

    def apply_distributed_callback(self, pattern, local_src, gather, dst, stream=None):
        """Callback overload: gather(needed_global_src_ids) -> values (host), called once."""
        err = []

        def cb(ids, n, out, user):
            try:
                v = np.asarray(gather(np.ctypeslib.as_array(ids, (n,)).copy() if n else np.zeros(0, np.int64)), dtype=np.float64)
                if v.shape != (n,):
                    err.append(f"gather returned {v.size} values but pattern.num_remote() is {n}")
                    return 1
                if n:
                    np.ctypeslib.as_array(out, (n,))[:] = v
                return 0
            except Exception as e:  # noqa: BLE001 -- reported after the call
                err.append(repr(e))
                return 2

        fn = GATHER_FN(cb)
        rc = lib().axis_apply_distributed_callback_c(self.handle, pattern.handle, _ptr(local_src), local_src.numel(), fn, None, _ptr(dst),
                                                     dst.numel(), _stream(stream))
        if err:
            raise AxisError("axis::solver::apply (callback): " + err[0])
        _check(rc)
        return dst

    def autotune(self, src, dst, n_fields, src_dtype, dst_dtype, mode, stream=None):
        """Times the launch shapes on these fields; returns (fields per step, field chunks, staging:
        0 direct gathers, 1 workgroup tiles, 2 wave tiles)."""
        t, r, st = C.c_int(0), C.c_int(0), C.c_int(0)
        _check(lib().axis_autotune_apply_c(self.handle, _ptr(src), src_dtype, _ptr(dst), dst_dtype, n_fields, mode,
                                           _stream(stream), C.byref(t), C.byref(r), C.byref(st)))
        return t.value, r.value, st.value

    def set_shape(self, n_fields, src_dtype, dst_dtype, mode, fields_per_step, field_chunks, staged):
        """Pins the launch shape of batch_apply for these field types and count
        (staged: 0 direct gathers, 1 workgroup tiles, 2 wave tiles)."""
        _check(lib().axis_apply_set_shape_c(self.handle, n_fields, src_dtype, dst_dtype, mode, fields_per_step, field_chunks, int(staged)))

    def check_conservation(self, src, dst, norm_type=0):
        out = np.zeros(4)
        if isinstance(src, np.ndarray):
            src = np.ascontiguousarray(src, dtype=np.float64)
        if isinstance(dst, np.ndarray):
            dst = np.ascontiguousarray(dst, dtype=np.float64)
        _check(lib().axis_check_conservation_c(self.handle, _ptr(src), _ptr(dst), norm_type, out.ctypes.data_as(_D)))
        return {"src_integral": out[0], "dst_integral": out[1], "absolute_error": out[2], "relative_error": out[3]}

    def close(self):
        if getattr(self, "handle", 0):
            _check(lib().axis_destroy_matrix_c(self.handle))
            self.handle = 0

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def generate_weights(src: Mesh, dst: Mesh, method: int = CONSERVATIVE) -> Matrix:
    """axis_generate_weights_c: method 0 Bilinear, 2 Conservative1stOrder."""
    h = C.c_int(0)
    _check(lib().axis_generate_weights_c(src.handle, dst.handle, method, C.byref(h)))
    return Matrix(h.value)


def generate(src: Mesh, dst: Mesh, config: RegridConfig) -> Matrix:
    h = C.c_int(0)
    _check(lib().axis_generate_weights_config_c(src.handle, dst.handle, C.byref(config), C.byref(h)))
    return Matrix(h.value)


class HaloPattern:
    """Off-rank source gather plan of one rank (HaloPattern, halo_pattern.hpp:29-48)."""

    def __init__(self, handle: int):
        self.handle = handle

    def info(self):
        nr, nk = C.c_int64(0), C.c_int64(0)
        _check(lib().axis_halo_pattern_info_c(self.handle, C.byref(nr), C.byref(nk)))
        return nr.value, nk.value

    def arrays(self):
        """dict of source_ranks, rank_offsets, needed_global_src_ids, gather_slot."""
        nr, nk = self.info()
        out = {"source_ranks": np.zeros(nk, np.int32), "rank_offsets": np.zeros(nk + 1, np.int64),
               "needed_global_src_ids": np.zeros(nr, np.int64), "gather_slot": np.zeros(nr, np.int64)}
        _check(lib().axis_halo_pattern_arrays_c(self.handle, out["source_ranks"].ctypes.data_as(_I32),
                                                out["rank_offsets"].ctypes.data_as(_I64),
                                                out["needed_global_src_ids"].ctypes.data_as(_I64),
                                                out["gather_slot"].ctypes.data_as(_I64)))
        return out

    def close(self):
        if getattr(self, "handle", 0):
            _check(lib().axis_destroy_halo_pattern_c(self.handle))
            self.handle = 0


def generate_distributed(src: "Mesh", dst: "Mesh", config, src_global_ids, owner_of_src, dst_global_ids=None):
    """Distributed generate of one rank: (Matrix with off-rank columns remapped, HaloPattern)."""
    sg = np.ascontiguousarray(src_global_ids, dtype=np.int64)
    ow = np.ascontiguousarray(owner_of_src, dtype=np.int32)
    dg = None if dst_global_ids is None else np.ascontiguousarray(dst_global_ids, dtype=np.int64)
    m, p = C.c_int(0), C.c_int(0)
    _check(lib().axis_generate_weights_distributed_c(src.handle, dst.handle, None if config is None else C.byref(config),
                                                     sg.ctypes.data_as(_I64), sg.shape[0],
                                                     None if dg is None else dg.ctypes.data_as(_I64), 0 if dg is None else dg.shape[0],
                                                     ow.ctypes.data_as(_I32), ow.shape[0], C.byref(m), C.byref(p)))
    return Matrix(m.value), HaloPattern(p.value)


def write_esmf(path, matrix: Matrix):
    """EsmfWeightIO::write_esmf: S, 1-based col_idx / row_idx, n_a / n_b (NetCDF classic file)."""
    _check(lib().axis_matrix_write_esmf_c(matrix.handle, str(path).encode()))


def read_esmf(path) -> Matrix:
    """EsmfWeightIO::read_esmf (NetCDF classic files; fractions and areas come back as zeros)."""
    h = C.c_int(0)
    _check(lib().axis_matrix_read_esmf_c(str(path).encode(), C.byref(h)))
    return Matrix(h.value)


def matrix_from_coo(rows, cols, vals, n_src, n_dst, frac_a=None, frac_b=None, area_a=None, area_b=None) -> Matrix:
    rows = np.ascontiguousarray(rows, dtype=np.int64)
    cols = np.ascontiguousarray(cols, dtype=np.int64)
    vals = np.ascontiguousarray(vals, dtype=np.float64)

    def opt(a):
        return None if a is None else np.ascontiguousarray(a, dtype=np.float64)

    fa, fb, aa, ab = opt(frac_a), opt(frac_b), opt(area_a), opt(area_b)
    h = C.c_int(0)
    _check(lib().axis_matrix_from_coo_c(rows.ctypes.data_as(_I64), cols.ctypes.data_as(_I64), vals.ctypes.data_as(_D),
                                        rows.shape[0], n_src, n_dst,
                                        *(None if a is None else a.ctypes.data_as(_D) for a in (fa, fb, aa, ab)),
                                        C.byref(h)))
    return Matrix(h.value)


class BatchPlan:
    """Several (matrix, src, dst, k) regrid pairs applied by one HIP graph
    (axis_batch_plan_*_c): a persistent work-queue launch, or -- when the
    autotuner measures it faster -- the pairs' own tuned launches captured in
    order.  The field tensors are bound at construction and must stay alive;
    run() replays the graph."""

    def __init__(self, pairs, src_dtype, dst_dtype, mode, autotune=False, stream=None, pair_launches=False):
        n = len(pairs)
        hs = (C.c_int * n)(*[m.handle for m, _, _, _ in pairs])
        xs = (C.c_void_p * n)(*[_ptr(x) for _, x, _, _ in pairs])
        ys = (C.c_void_p * n)(*[_ptr(y) for _, _, y, _ in pairs])
        ks = (C.c_int64 * n)(*[int(k) for _, _, _, k in pairs])
        self._keep = list(pairs)
        h = C.c_int(0)
        _check(lib().axis_batch_plan_create_c(n, hs, xs, ys, ks, src_dtype, dst_dtype, mode, 2 if pair_launches else (1 if autotune else 0),
                                              _stream(stream), C.byref(h)))
        self.handle = h.value

    def run(self, stream=None, sync=True):
        _check(lib().axis_batch_plan_run_c(self.handle, _stream(stream), 1 if sync else 0))

    def info(self):
        fs, fi, nw = C.c_int(0), C.c_int(0), C.c_int(0)
        ni = C.c_int64(0)
        _check(lib().axis_batch_plan_info_c(self.handle, C.byref(fs), C.byref(fi), C.byref(ni), C.byref(nw)))
        rec = "pair_launches" if fs.value == 0 else "persistent"
        return {"recording": rec, "fields_per_step": fs.value, "fields_per_item": fi.value, "items": ni.value, "workgroups": nw.value}

    def close(self):
        if getattr(self, "handle", 0):
            lib().axis_destroy_batch_plan_c(self.handle)
            self.handle = 0

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


N_MLP_PARAMS = 17924


class MLP:
    """The candidate-proposer network on the GPU (axis_mlp_*_c): 6 -> 128 ->
    128 -> 4, BF16 MFMA with FP32 accumulation, Adam training."""

    def __init__(self, params=None, seed=0, feature_scale=1.0):
        h = C.c_int(0)
        p = None
        if params is not None:
            self._init = np.ascontiguousarray(params, dtype=np.float32)
            p = self._init.ctypes.data_as(C.POINTER(C.c_float))
        _check(lib().axis_mlp_create_c(p, C.c_uint64(seed), C.c_float(feature_scale), C.byref(h)))
        self.handle = h.value

    def params(self):
        out = np.zeros(N_MLP_PARAMS, dtype=np.float32)
        fs = C.c_float(0)
        _check(lib().axis_mlp_params_c(self.handle, out.ctypes.data_as(C.POINTER(C.c_float)), C.byref(fs)))
        return out

    def feature_scale(self):
        fs = C.c_float(0)
        _check(lib().axis_mlp_params_c(self.handle, None, C.byref(fs)))
        return fs.value

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        _check(lib().axis_mlp_set_params_c(self.handle, p.ctypes.data_as(C.POINTER(C.c_float))))

    def forward(self, x, out, stream=None):
        _check(lib().axis_mlp_forward_c(self.handle, _ptr(x), _ptr(out), int(x.shape[0]), _stream(stream)))
        return out

    def gradient(self, x, t, grad, wc=1.0, tau=0.5, stream=None):
        loss = C.c_float(0)
        _check(lib().axis_mlp_gradient_c(self.handle, _ptr(x), _ptr(t), int(x.shape[0]), C.c_float(wc), C.c_float(tau),
                                         _ptr(grad), C.byref(loss), _stream(stream)))
        return loss.value

    def train(self, x, t, batch, lr, wc=1.0, tau=0.5, stream=None):
        loss = C.c_float(0)
        _check(lib().axis_mlp_train_c(self.handle, _ptr(x), _ptr(t), int(x.shape[0]), int(batch), C.c_float(lr),
                                      C.c_float(wc), C.c_float(tau), C.byref(loss), _stream(stream)))
        return loss.value

    def close(self):
        if getattr(self, "handle", 0):
            lib().axis_destroy_mlp_c(self.handle)
            self.handle = 0

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def generate_weights_proposed(src: "Mesh", dst: "Mesh", net: MLP, config: "RegridConfig" = None, area_tol: float = 1e-12):
    """Conservative weights whose broad phase uses the proposer network
    (axis_generate_weights_proposed_c); returns (Matrix, stats dict)."""
    h = C.c_int(0)
    st = np.zeros(5, dtype=np.int64)
    cfg = C.byref(config) if config is not None else None
    _check(lib().axis_generate_weights_proposed_c(src.handle, dst.handle, cfg, net.handle, C.c_double(area_tol), C.byref(h),
                                                  st.ctypes.data_as(C.POINTER(C.c_int64))))
    keys = ("cells", "guard_rejected", "area_failed", "pairs_proposed", "pairs_fallback")
    return Matrix(h.value), dict(zip(keys, (int(v) for v in st)))


def proposer_training_data(src: "Mesh", dst: "Mesh", feature_scale: float, x, t):
    """Fills device tensors x [n_dst, 6], t [n_dst, 4] with the proposer's
    features and labels for this mesh pair; returns the unlabelled count."""
    bad = C.c_int64(0)
    _check(lib().axis_proposer_training_data_c(src.handle, dst.handle, C.c_float(feature_scale), _ptr(x), _ptr(t),
                                               C.byref(bad)))
    return bad.value
