/* axis_hip.h -- C ABI of the MI355X-native AXIS regrid library (libaxis_hip.so).
 *
 * Drop-in for the reference's flat C interface
 * (libs/axis/src/fortran/axis_c_interop.cpp, bound from Fortran by
 * libs/axis/fortran/axis_mod.f90): integer handles (token 0 invalid,
 * monotonic, never recycled), AXIS_SUCCESS / AXIS_ERROR return codes, no
 * exception crosses the boundary.  Calls are synchronous unless they take a
 * stream.  Every entry point below cites the reference interface it replaces;
 * entry points without a reference counterpart say so.
 *
 * Meshes and weight matrices live in GPU memory behind their handles.  Field
 * pointers passed to the apply entry points may be host or device memory where
 * stated.
 */
#ifndef AXIS_HIP_H
#define AXIS_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define AXIS_SUCCESS 0 /* axis_c_interop.cpp:37 */
#define AXIS_ERROR -1  /* axis_c_interop.cpp:40 */

/* InterpolationMethod codes of axis_generate_weights_c (axis_c_interop.cpp:126-140) */
#define AXIS_METHOD_BILINEAR 0
#define AXIS_METHOD_NEAREST 1
#define AXIS_METHOD_CONSERVATIVE1 2

/* CoordinateSystem (libs/axis/include/axis/topology/enums.hpp:73-80) */
#define AXIS_COORD_SPHERICAL_DEG 0
#define AXIS_COORD_SPHERICAL_RAD 1
#define AXIS_COORD_CARTESIAN3D 2

/* field storage types and apply modes (no reference counterpart: the
 * reference applies FP64 fields only) */
#define AXIS_F64 0
#define AXIS_F32 1
#define AXIS_BF16 2
#define AXIS_APPLY_PLAIN 0 /* FP64 fields, FP64 weights and accumulation */
#define AXIS_APPLY_MIXED 1 /* FP32/BF16 fields, FP32 hi+lo weights, compensated FP32 accumulation */
#define AXIS_APPLY_MIXED_ACC64 2 /* FP32/BF16 fields, FP64 weights, FP64 accumulation */

/* RegridConfig (libs/axis/include/axis/solver/regrid_config.hpp:80-89);
 * field values are the enums' declaration order. */
typedef struct axis_regrid_config {
    int method;        /* InterpolationMethod: 0 Bilinear .. 4 Conservative1stOrder, 5 Conservative2ndOrder */
    int norm_type;     /* NormType: 0 DstArea, 1 FracArea */
    int line_type;     /* LineType: 0 Cartesian, 1 GreatCircle */
    int unmapped;      /* UnmappedAction: 0 Error, 1 Ignore */
    int use_limiter;   /* unused by first-order methods */
    int extrap_method; /* ExtrapolationAction: 0 None, 1 NearestWet */
} axis_regrid_config;

/* axis_init_c (axis_c_interop.cpp:88-90): idempotent runtime initialisation. */
int axis_init_c(void);

/* Message of the last AXIS_ERROR returned on this thread (no reference
 * counterpart: the reference drops the exception text at the boundary). */
const char *axis_last_error_c(void);

/* Mesh from UGRID-style buffers: node_coords [n_nodes, ndim] column-major
 * (all lon, then all lat), CSR cell->node connectivity, optional cell areas
 * and mask (NULL when absent).  Replaces axis_mesh_from_descriptor_c
 * (axis_c_interop.cpp:105-109) for the UGRID descriptor path
 * (mesh_factory.cpp:314-339 build_from_ugrid); the arrays are copied. */
int axis_mesh_create_c(const double *node_coords, int64_t n_nodes, int64_t ndim, const int64_t *conn_offsets, int64_t n_cells,
                       const int64_t *conn_indices, int coord_sys, const double *cell_areas, const int32_t *cell_mask, int *mesh_handle);

/* Regular lat-lon mesh by rules: RuleGenerator RegularLatLon
 * (rule_generator.cpp:105-169, reached through MeshFactory::from_rules). */
int axis_mesh_regular_latlon_c(double min_x, double max_x, double min_y, double max_y, double r_x, double r_y, int *mesh_handle);

/* axis_mesh_from_descriptor_c (axis_c_interop.cpp:105-109): descriptor_ptr is
 * the address of a C++ axis::ingest::GridDescriptor
 * (ingest/grid_descriptor.hpp:192-216) filled by the caller and read in place
 * (MeshFactory::from_descriptor, mesh_factory.cpp:396-431, with its validation
 * messages).  Served kinds: UGRID, CF and GRIB (StructuredGrid::to_unstructured,
 * corners synthesized from the centres when absent) and GridRules of kind
 * RegularLatLon; Projected, NamedGrid and the other rule kinds return
 * AXIS_ERROR with a message. */
int axis_mesh_from_descriptor_c(const void *descriptor_ptr, int *mesh_handle);

/* axis_mesh_from_named_c (axis_c_interop.cpp:96-99): the named-grid registry
 * (named_grid_registry.cpp) is not part of this library; always AXIS_ERROR
 * with the message "named grids are not built". */
int axis_mesh_from_named_c(const char *name, int *mesh_handle);

/* node_coords() / conn_offsets() / conn_indices() of UnstructuredMesh
 * (unstructured_mesh.hpp:113-137): copies into host buffers sized from
 * axis_mesh_info_c (node_coords 2 * n_nodes doubles, all lon then all lat;
 * conn_offsets n_cells + 1; conn_indices conn_offsets[n_cells]); any pointer
 * may be NULL. */
int axis_mesh_arrays_c(int mesh_handle, double *node_coords, int64_t *conn_offsets, int64_t *conn_indices);

/* Sizes of a mesh (UnstructuredMesh::n_nodes / n_cells, unstructured_mesh.hpp:86-97). */
int axis_mesh_info_c(int mesh_handle, int64_t *n_nodes, int64_t *n_cells);

/* Grid structure as the weight generators see it: *kind = 0 unstructured,
 * 1 rectilinear (not regular), 2 regular lat-lon; ninj[2] = (ni, nj) for the
 * last two; out6 = (lon_min, lon_max, delta_lon, lat_min, lat_max, delta_lat)
 * for a regular grid.  Replaces detect_regular_grid / detect_rectilinear_grid
 * (detail/regular_grid_detector.hpp:64-144, 180-299). */
int axis_mesh_grid_info_c(int mesh_handle, int *kind, int64_t *ninj, double *out6);

/* ---- distributed mode of one rank (the exchange itself is the caller's) */

/* HaloPattern handle: the off-rank source cells a rank's apply reads,
 * grouped by owning rank (halo_pattern.hpp:29-48).  Produced by
 * WeightGenerator::generate(src, dst, config, src_global_ids, dst_global_ids,
 * owner_of_src) (weight_generator.hpp:83-90, weight_generator.cpp:4525-4680):
 * the local matrix's off-rank columns are remapped to n_local + slot.  The
 * rank is the owner of the first local source cell; config may be NULL
 * (defaults).  dst_global_ids is accepted and, as in the reference, unused. */
int axis_generate_weights_distributed_c(int src_handle, int dst_handle, const axis_regrid_config *config, const int64_t *src_global_ids,
                                        int64_t n_src_ids, const int64_t *dst_global_ids, int64_t n_dst_ids, const int32_t *owner_of_src,
                                        int64_t n_owner, int *matrix_handle, int *pattern_handle);
/* HaloPattern::num_remote / num_ranks (halo_pattern.hpp:39-47). */
int axis_halo_pattern_info_c(int pattern_handle, int64_t *num_remote, int64_t *num_ranks);
/* Copies of the pattern's arrays into host buffers sized from
 * axis_halo_pattern_info_c (rank_offsets has num_ranks + 1 entries); NULL
 * skips an array. */
int axis_halo_pattern_arrays_c(int pattern_handle, int32_t *source_ranks, int64_t *rank_offsets, int64_t *needed_global_src_ids,
                               int64_t *gather_slot);
int axis_destroy_halo_pattern_c(int pattern_handle);
/* Distributed apply with the gathered halo buffer (apply.hpp:151-211):
 * dst = S [local_src ; gathered_halo_src] on FP64 device fields, layout_left
 * [n, n_fields]; the reference's extent checks and messages, before any
 * write. */
int axis_apply_distributed_c(int matrix_handle, int pattern_handle, const double *local_src, int64_t n_local, const double *gathered_halo_src,
                             int64_t n_halo, double *dst, int64_t n_dst, int64_t n_fields, void *stream);
/* Gather callback of the callback overload: fills out[n_remote] (host) with
 * the values of needed_global_src_ids in gather_slot order; nonzero = error. */
typedef int (*axis_gather_fn)(const int64_t *needed_global_src_ids, int64_t n_remote, double *out, void *user);
/* Callback distributed apply (apply.hpp:233-248): gather is invoked exactly
 * once, then the buffer overload runs (one FP64 device field). */
int axis_apply_distributed_callback_c(int matrix_handle, int pattern_handle, const double *local_src, int64_t n_local, axis_gather_fn gather,
                                      void *user, double *dst, int64_t n_dst, void *stream);

/* Cell centroids as the masked-source extrapolation uses them:
 * compute_cell_centroids_xy (weight_generator.cpp:74-138) rounded to float
 * (:1813, :1843), into host arrays lon[n_cells], lat[n_cells].  Computed on the
 * GPU; the cells whose float could differ from host libm's are recomputed on
 * the host, and *n_host (optional) receives their count.  Diagnostic: the
 * reference computes these inside coastal_renormalize_and_extrapolate only. */
int axis_mesh_cell_centroids_c(int mesh_handle, float *lon, float *lat, int64_t *n_host);

/* axis_generate_weights_c (axis_c_interop.cpp:117-145): method 0 Bilinear,
 * 2 Conservative1stOrder with the default RegridConfig otherwise. */
int axis_generate_weights_c(int src_handle, int dst_handle, int method, int *matrix_handle);

/* axis_generate_vector_weights_c (axis_c_interop.cpp:156-205,
 * VectorWeightGenerator, vector_regridder.cpp:14): rotated U/V weights are
 * not part of this library; the handles are checked, then AXIS_ERROR with an
 * explicit message and the output handles untouched. */
int axis_generate_vector_weights_c(int src_handle, int dst_handle, const double *src_alpha, const double *dst_alpha, int method, int *matrix_u_handle,
                                   int *matrix_v_handle);

/* WeightGenerator::generate with a full RegridConfig (weight_generator.hpp:60-62). */
int axis_generate_weights_config_c(int src_handle, int dst_handle, const axis_regrid_config *config, int *matrix_handle);

/* InterpolationMatrix from COO triplets (interpolation_matrix.hpp:80-92);
 * the matrix is stored in to_csr() order, duplicates kept.  frac/area arrays
 * may be NULL (filled with 1). */
int axis_matrix_from_coo_c(const int64_t *rows, const int64_t *cols, const double *values, int64_t nnz, int64_t n_src, int64_t n_dst,
                           const double *frac_a, const double *frac_b, const double *area_a, const double *area_b, int *matrix_handle);

/* InterpolationMatrix::to_csr (interpolation_matrix.hpp:180-255).  Matrices
 * here are always held in CSR form, so this only validates the handle. */
int axis_matrix_to_csr_c(int matrix_handle);

/* n_src(), n_dst(), nnz() (interpolation_matrix.hpp:98-112). */
int axis_matrix_info_c(int matrix_handle, int64_t *n_src, int64_t *n_dst, int64_t *nnz);

/* row_ptr() / col_idx() / csr_values() (interpolation_matrix.hpp:262-295):
 * copies into caller buffers (host or device) of n_dst+1, nnz, nnz entries. */
int axis_matrix_csr_c(int matrix_handle, int64_t *row_ptr, int64_t *col_idx, double *values);

/* frac_a() / frac_b() / area_a() / area_b() (interpolation_matrix.hpp:139-161);
 * any pointer may be NULL. */
int axis_matrix_fractions_c(int matrix_handle, double *frac_a, double *frac_b, double *area_a, double *area_b);

/* Generation diagnostics (no reference counterpart): out[0] candidate pairs,
 * out[1] pairs re-decided on the host, out[2] cap-test ties, out[3] longest
 * row; *gen_ms wall time of the generate call. */
int axis_matrix_stats_c(int matrix_handle, int64_t *out4, double *gen_ms);

/* Generation diagnostics (no reference counterpart): GPU time in ms, from HIP
 * events on the generation stream, of the general conservative path's phases
 * -- out[0] broad-phase query, out[1] cap filter kernel, out[2] clip kernel,
 * out[3] CSR assembly -- -1 for a phase the generation did not run (other
 * methods, analytic paths).  Fills out[0 .. n). */
int axis_matrix_phase_ms_c(int matrix_handle, double *out, int n);

/* axis_apply_c (axis_c_interop.cpp:217-236): dst = S * src, FP64.  src/dst may
 * be host or device memory. */
int axis_apply_c(int matrix_handle, const double *src, double *dst, int n_src, int n_dst);

/* batch_apply (apply.hpp:275-343, :336) on device-resident layout_left fields
 * src [n_src, n_fields], dst [n_dst, n_fields] of the given storage types,
 * enqueued on `stream` (a hipStream_t, NULL = default stream) and waited for
 * before returning.  mode AXIS_APPLY_PLAIN needs F64/F64; the mixed modes
 * take F32 or BF16 src and F32 or F64 dst. */
int axis_batch_apply_c(int matrix_handle, const void *src, int src_dtype, void *dst, int dst_dtype, int64_t n_src, int64_t n_dst,
                       int64_t n_fields, int mode, void *stream);

/* Same as axis_batch_apply_c but returns once the kernels are enqueued (no
 * reference counterpart: the reference apply always fences). */
int axis_batch_apply_async_c(int matrix_handle, const void *src, int src_dtype, void *dst, int dst_dtype, int64_t n_src, int64_t n_dst,
                             int64_t n_fields, int mode, void *stream);

/* Batched apply that allocates nothing (no counterpart in the reference; for
 * callers that cannot allocate device memory: code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  Same arguments and
 * results as axis_batch_apply_async_c (enqueue only: no synchronisation, no
 * memset), with two limits: the matrix's apply layout must already exist (it
 * is built by the first axis_batch_apply_c / axis_batch_apply_async_c /
 * axis_autotune_apply_c of the matrix with the same field types and count;
 * otherwise AXIS_ERROR), and the long-row partial sums (long-row segments x
 * n_fields doubles) must fit a fixed per-GPU array of 4,194,304 doubles
 * (32 MiB, held until the process ends; larger requests return AXIS_ERROR).
 * Every call on one GPU shares that array: calls that overlap on one GPU
 * race. */
int axis_batch_apply_noalloc_c(int matrix_handle, const void *src, int src_dtype, void *dst, int dst_dtype, int64_t n_src, int64_t n_dst,
                               int64_t n_fields, int mode, void *stream);

/* Autotuner (no reference counterpart): times the launch shapes of
 * axis_batch_apply_c (fields per step x field chunks x direct gathers,
 * LDS-staged workgroup tiles or LDS-staged wave tiles) on these device
 * fields, caches the fastest for this (matrix size, n_fields, types, mode)
 * and reports it (*staged: 0 direct, 1 workgroup tiles, 2 wave tiles);
 * output pointers may be NULL. */
int axis_autotune_apply_c(int matrix_handle, const void *src, int src_dtype, void *dst, int dst_dtype, int64_t n_fields, int mode, void *stream,
                          int *fields_per_step, int *field_chunks, int *staged);

/* Pins the launch shape of axis_batch_apply_c for this (matrix size,
 * n_fields, types, mode) as the autotuner's cache entry would (no reference
 * counterpart): fields per step 1, 2 or 4, field chunks >= 1, staged 0
 * direct gathers, 1 workgroup tiles, 2 wave tiles.  Every shape gives the
 * same bits. */
int axis_apply_set_shape_c(int matrix_handle, int64_t n_fields, int src_dtype, int dst_dtype, int mode, int fields_per_step, int field_chunks,
                           int staged);

/* check_conservation (conservation.hpp:95-97, conservation.cpp:82-103) on
 * FP64 fields (host or device): out4 = {src_integral, dst_integral,
 * absolute_error, relative_error}. */
int axis_check_conservation_c(int matrix_handle, const double *src, const double *dst, int norm_type, double *out4);

/* Multi-pair persistent batch (no reference counterpart; the reference
 * applies one matrix per batch_apply call, apply.hpp:275-343).  Binds
 * n_pairs (matrix, device src [n_src, n_fields[p]], device dst
 * [n_dst, n_fields[p]]) triples of one field type and mode, builds the work
 * queues and records one HIP graph.  autotune 0: the persistent work-queue
 * kernel (+ long-row reductions) with a default shape; 1: its shapes and the
 * pairs' own autotuned launches captured in order are timed and the faster
 * recording kept; 2: the pairs' own launches (tuned shapes).  The matrices
 * stay alive while the plan exists. */
int axis_batch_plan_create_c(int n_pairs, const int *matrix_handles, const void *const *src, void *const *dst, const int64_t *n_fields,
                             int src_dtype, int dst_dtype, int mode, int autotune, void *stream, int *plan_handle);

/* Replays the plan's graph on `stream` (NULL = default stream); sync != 0
 * waits for it.  Replays of one plan must not overlap. */
int axis_batch_plan_run_c(int plan_handle, void *stream, int sync);

/* Chosen shape (fields per step, fields per work item), work items and
 * persistent workgroups of a plan; pointers may be NULL. */
int axis_batch_plan_info_c(int plan_handle, int *fields_per_step, int *fields_per_item, int64_t *n_items, int *n_workgroups);

/* Releases a plan handle (the graph and the plan's workspaces). */
int axis_destroy_batch_plan_c(int handle);

/* Candidate-proposer network (no reference counterpart; the reference's
 * broad phase is an ArborX BVH only).  A 6 -> 128 -> 128 -> 4 MLP with BF16
 * weights on the MFMA units and FP32 accumulation: inputs are a destination
 * cell's centroid xyz and its box extents / feature_scale, outputs the
 * proposed search-cap centre offset (units of feature_scale) and log of its
 * radius (units of feature_scale).  params is NULL (seeded initialisation)
 * or 17,924 FP32 values (W1 [128][6], b1, W2 [128][128], b2, W3 [4][128], b3). */
int axis_mlp_create_c(const float *params, uint64_t seed, float feature_scale, int *mlp_handle);
/* Copies the FP32 parameters (17,924 values) and/or the feature scale out; either pointer may be NULL. */
int axis_mlp_params_c(int mlp_handle, float *params, float *feature_scale);
int axis_mlp_set_params_c(int mlp_handle, const float *params);
/* out [n][4] = network(x [n][6]); device pointers. */
int axis_mlp_forward_c(int mlp_handle, const float *x, float *out, int64_t n, void *stream);
/* Mean loss over (x [n][6], t [n][4]) and its gradient into grad [17,924]
 * (device pointers): wc * |o[0:3] - t[0:3]|^2 + pinball_tau(t[3] - o[3]). */
int axis_mlp_gradient_c(int mlp_handle, const float *x, const float *t, int64_t n, float wc, float tau, float *grad, float *loss, void *stream);
/* One Adam pass over the samples in order, mini-batches of `batch`; *mean_loss = mean batch loss. */
int axis_mlp_train_c(int mlp_handle, const float *x, const float *t, int64_t n, int64_t batch, float lr, float wc, float tau, float *mean_loss,
                     void *stream);
int axis_destroy_mlp_c(int handle);

/* First-order conservative weights (config may be NULL: the default
 * RegridConfig with method Conservative1stOrder) whose broad phase takes the
 * proposer network's search caps (no reference counterpart).  A proposal is
 * used only when it provably contains every source cell the plain method
 * could clip against the cell and the clipped overlaps sum to the cell's area
 * within area_tol relative (<= 0: 1e-12); other cells go through the LBVH,
 * so the matrix equals axis_generate_weights_config_c's.  stats5 (may be
 * NULL) = {cells, guard rejections, area-check failures, proposed candidate
 * pairs, LBVH candidate pairs of the redone cells}. */
int axis_generate_weights_proposed_c(int src_handle, int dst_handle, const axis_regrid_config *config, int mlp_handle, double area_tol,
                                     int *matrix_handle, int64_t *stats5);

/* Training pairs for the proposer from one (source, destination) mesh pair:
 * x [n_dst][6] features, t [n_dst][4] labels (zero centre offset, log of the
 * containment guard radius / feature_scale); device pointers. */
int axis_proposer_training_data_c(int src_handle, int dst_handle, float feature_scale, float *x, float *t, int64_t *n_unlabelled);

/* nan_batch_apply (apply.hpp:359-420): FP64 layout_left fields [n, n_fields]
 * as device pointers; NaN sources are skipped and each row renormalised by
 * its valid weight; the output is NaN where the valid weight fraction is
 * below 1 - na_thres - 1e-6 (or zero).  Synchronous on `stream`. */
int axis_nan_batch_apply_c(int matrix_handle, const double *src, double *dst, int64_t n_src, int64_t n_dst, int64_t n_fields, double na_thres,
                           void *stream);

/* Masked apply (apply.hpp:438-484, the dst_mask overload of apply): rows with
 * dst_mask == 0 keep their values; the mask applies when n_mask == n_dst.
 * FP64 layout_left fields (device pointers), n_fields >= 1.  Synchronous. */
int axis_apply_masked_c(int matrix_handle, const double *src, double *dst, const int32_t *dst_mask, int64_t n_mask, int64_t n_src, int64_t n_dst,
                        int64_t n_fields, void *stream);

/* Weight blob in the AXWC format (WeightCache::serialize / deserialize,
 * weight_cache.hpp:64-157, 286-296): 40-byte header, then factor_list,
 * factor_row, factor_col, frac_a, frac_b, area_a, area_b.  buf == NULL asks
 * for the size (*bytes); the caller owns the buffer.  Deserialize validates
 * magic, version, endianness and size with the reference's messages. */
int axis_matrix_serialize_c(int matrix_handle, void *buf, uint64_t buf_size, uint64_t *bytes);
int axis_matrix_deserialize_c(const void *buf, uint64_t buf_size, int *matrix_handle);

/* ESMF weight files (EsmfWeightIO::write_esmf / read_esmf,
 * esmf_weight_io.cpp:68-173): num_wgts, S, 1-based col_idx / row_idx, global
 * n_a / n_b.  Written in the NetCDF classic 64-bit-offset format; the
 * classic formats are read, a NetCDF-4 (HDF5) file is refused. */
int axis_matrix_write_esmf_c(int matrix_handle, const char *path);
int axis_matrix_read_esmf_c(const char *path, int *matrix_handle);

/* axis_destroy_mesh_c / axis_destroy_matrix_c (axis_c_interop.cpp:241-250). */
int axis_destroy_mesh_c(int handle);
int axis_destroy_matrix_c(int handle);

#ifdef __cplusplus
}
#endif

#endif /* AXIS_HIP_H */
