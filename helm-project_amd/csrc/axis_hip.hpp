// axis_hip.hpp -- C++20 host API of the MI355X AXIS regrid library.
//
// Mirrors the reference's solver API (libs/axis):
//   WeightGenerator::generate(src, dst, RegridConfig) -> InterpolationMatrix
//                                           weight_generator.hpp:60-62
//   InterpolationMatrix (COO views, CSR row_ptr/col_idx/values, frac/area)
//                                           interpolation_matrix.hpp:61-372
//   apply / batch_apply (layout_left fields) apply.hpp:61-119, 275-343
//   check_conservation                      conservation.hpp:95-97
// with the same argument meaning and error behaviour (InvalidArgument thrown
// before any output is written, RuntimeError for unmapped cells / HIP
// failures).  Meshes and matrices live in HBM; host arrays are copied in once
// at construction.
#pragma once

#include <cstddef>
#include <cstdint>
#include <atomic>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "common.hpp"

namespace axis_hip {

namespace mlp {
class Model;
}
struct ProposerStats;

// regrid_config.hpp:24-89
enum class InterpolationMethod : std::uint8_t { Bilinear, NearestNeighbor, Bicubic, Patch, Conservative1stOrder, Conservative2ndOrder };
enum class NormType : std::uint8_t { DstArea, FracArea };
enum class LineType : std::uint8_t { Cartesian, GreatCircle };
enum class UnmappedAction : std::uint8_t { Error, Ignore };
enum class ExtrapolationAction : std::uint8_t { None, NearestWet };
// enums.hpp:73-80
enum class CoordinateSystem : std::uint8_t { SphericalDeg, SphericalRad, Cartesian3D };

struct RegridConfig {
    InterpolationMethod method = InterpolationMethod::Bilinear;
    NormType norm_type = NormType::DstArea;
    LineType line_type = LineType::GreatCircle;
    UnmappedAction unmapped = UnmappedAction::Ignore;
    bool use_limiter = false;
    ExtrapolationAction extrap_method = ExtrapolationAction::NearestWet;
};

/// UnstructuredMesh (unstructured_mesh.hpp:41-220): node_coords [n_nodes, 2]
/// layout_left (lon block, then lat block), CSR cell->node connectivity,
/// optional per-cell areas and mask.  A host copy is kept for the few
/// libm-exact host steps; the device copy feeds the kernels.
class Mesh {
   public:
    Mesh(const double *node_coords, index_t n_nodes, index_t ndim, const index_t *conn_offsets, index_t n_cells, const index_t *conn_indices,
         CoordinateSystem cs, const double *areas, const std::int32_t *mask);
    index_t n_nodes() const { return n_nodes_; }
    index_t n_cells() const { return n_cells_; }
    CoordinateSystem coord_system() const { return cs_; }
    bool has_areas() const { return !h_areas.empty(); }
    bool has_mask() const { return !h_mask.empty(); }
    int max_cell_vertices() const { return max_nv_; }
    int min_cell_vertices() const { return min_nv_; }

    // host copies
    std::vector<double> h_coords;  // [2 * n_nodes] lon block then lat block
    std::vector<index_t> h_offsets, h_indices;
    std::vector<double> h_areas;
    std::vector<std::int32_t> h_mask;
    // device copies
    DeviceBuffer<double> d_coords;
    DeviceBuffer<index_t> d_offsets, d_indices;
    DeviceBuffer<double> d_areas;
    DeviceBuffer<std::int32_t> d_mask;

   private:
    index_t n_nodes_ = 0, n_cells_ = 0;
    CoordinateSystem cs_ = CoordinateSystem::SphericalDeg;
    int max_nv_ = 0, min_nv_ = 0;
};

/// InterpolationMatrix, always held in CSR form sorted by (row, col)
/// (the layout to_csr() produces, interpolation_matrix.hpp:180-255), in HBM.
class InterpolationMatrix {
   public:
    index_t n_src = 0, n_dst = 0, nnz = 0;
    DeviceBuffer<index_t> row_ptr;   // [n_dst + 1]
    DeviceBuffer<index_t> col;       // [nnz]
    DeviceBuffer<double> val;        // [nnz]
    DeviceBuffer<std::int32_t> col32;  // [nnz] int32 column copy streamed by the apply kernels
    DeviceBuffer<float> val_hi, val_lo;  // [nnz] FP64 weights split into FP32 hi + lo
    DeviceBuffer<double> frac_a, area_a;  // [n_src]
    DeviceBuffer<double> frac_b, area_b;  // [n_dst]
    int max_row_nnz = 0;
    // Apply-side layout, built once by prepare_apply(): rows cut into
    // segments of at most seg_cap (E) nonzeros, in processing order.  A row
    // longer than E (a "long row") owns slots long_slot[j] .. long_slot[j+1]
    // of the FP64 partial-sum workspace, one per segment.
    int seg_cap = 16;
    index_t n_seg = 0;
    DeviceBuffer<std::int32_t> seg_row, seg_len, seg_slot;
    DeviceBuffer<index_t> seg_begin;
    int n_long_rows = 0, n_slots = 0;
    DeviceBuffer<std::int32_t> long_rows, long_slot;
    // Long-row partial-sum workspaces, one per stream the matrix is applied
    // on: applies on one stream are ordered by it, applies on different
    // streams write different workspaces, so concurrent applies of one matrix
    // do not race.  apply_mu guards them and the lazily built layouts.
    std::map<hipStream_t, DeviceBuffer<double>> partial_ws;
    std::mutex apply_mu;
    // Spatial processing order: 30-bit Morton code of each destination cell's
    // centroid (unit-sphere xyz), set by the weight generators; empty for
    // matrices built from triplets (rows then keep index order).
    DeviceBuffer<std::uint32_t> row_key;
    // Tiles: tile t = segments [256 t, 256 t + 256).  A tile whose nonzeros
    // touch at most 256 distinct aligned 4-column chunks lists their first
    // columns (sorted) in tile_col[t * 256 ...], and t_lcol[e] (aligned with
    // the CSR arrays) holds nonzero e's position in the staged values
    // (4 * chunk position + column % 4); the tiled apply stages those chunks
    // in LDS once per field with 16-byte loads.  tile_nu[t] = -1 marks a tile
    // that gathers straight from HBM.
    index_t n_tiles = 0;
    DeviceBuffer<std::int32_t> tile_nu, tile_col;
    DeviceBuffer<std::uint16_t> t_lcol;
    bool apply_ready = false, tiles_ready = false;
    // Wave tiles: wave tile w = segments [64 w, 64 w + 64); per chunk width
    // (16-byte chunks of 2, 4 or 8 columns: FP64, FP32, BF16 fields) the
    // distinct chunks each wave stages (at most 128; wnu[w] = -1: too many,
    // the wave gathers directly), their first columns wcol[w * 128 ...], and
    // per nonzero its column's position in the staged chunks.
    index_t n_wtiles = 0;
    DeviceBuffer<std::int32_t> wnu[3], wcol[3];
    DeviceBuffer<std::uint16_t> wlpos[3];
    bool wave_ready[3] = {false, false, false};
    // row-length histogram (lengths 0 .. 16, then "longer"), cached by
    // row_length_hist(); it fixes seg_cap
    index_t len_hist[18] = {0};
    bool len_hist_ready = false;
    // diagnostics of the generation that produced it
    index_t n_candidates = 0, n_host_arbitrated = 0, n_cap_borderline = 0;
    double gen_ms = 0.0;
    // GPU time of the general conservative path's phases (HIP events on the
    // generation stream): broad-phase query, cap filter kernel, clip kernel,
    // CSR assembly; -1 where the path did not run them
    double phase_ms[4] = {-1.0, -1.0, -1.0, -1.0};

    /// Builds the apply layout once: FP32 hi/lo split, int32 columns and the
    /// segment list in spatial order.
    void prepare_apply(hipStream_t s);
    /// Builds the tiles once (on the first tiled launch).
    void prepare_tiles(hipStream_t s);
    /// Builds the wave tiles for chunks of cw columns once (first wave launch
    /// with fields of that width).
    void prepare_wave_tiles(int cw, hipStream_t s);
    /// Forgets every apply-side layout (after the CSR arrays change).
    void reset_apply_layout() {
        apply_ready = tiles_ready = false;
        for (bool &w : wave_ready) w = false;
    }
    /// Fills len_hist once; returns the segment length E it implies (the
    /// smallest of 4, 8, 12, 16 leaving at most 0.1% of the rows longer).
    int row_length_hist(hipStream_t s);
};

/// HaloPattern (halo_pattern.hpp:29-48): the off-rank source cells a rank's
/// apply reads, grouped by owning rank in CSR form: the ids needed from
/// source_ranks[r] are needed_global_src_ids[rank_offsets[r] ..
/// rank_offsets[r + 1]), and gather_slot[i] is id i's slot in the gathered
/// halo buffer.
struct HaloPattern {
    std::vector<int> source_ranks;
    std::vector<index_t> rank_offsets{0};
    std::vector<index_t> needed_global_src_ids;
    std::vector<index_t> gather_slot;
    index_t num_remote() const { return static_cast<index_t>(needed_global_src_ids.size()); }
    index_t num_ranks() const { return static_cast<index_t>(source_ranks.size()); }
};

struct ConservationReport {
    double src_integral = 0, dst_integral = 0, absolute_error = 0, relative_error = 0;
};

enum class FieldType : int { F64 = 0, F32 = 1, BF16 = 2 };
/// Plain: FP64 fields, weights and accumulation.  Mixed: FP32/BF16 fields,
/// FP32 hi + lo weights, TwoSum-compensated FP32 accumulation.  MixedAcc64:
/// FP32/BF16 fields, FP64 weights and FP64 accumulation.
enum class ApplyMode : int { Plain = 0, Mixed = 1, MixedAcc64 = 2 };

class WeightGenerator {
   public:
    static std::unique_ptr<InterpolationMatrix> generate(const Mesh &src, const Mesh &dst, const RegridConfig &cfg, hipStream_t s = nullptr);
    /// generate() whose conservative broad phase takes the proposer
    /// network's guarded search caps (LBVH for the cells they cannot serve);
    /// the weights equal generate()'s.
    static std::unique_ptr<InterpolationMatrix> generate_proposed(const Mesh &src, const Mesh &dst, const RegridConfig &cfg,
                                                                  const mlp::Model *proposer, ProposerStats *stats, hipStream_t s = nullptr,
                                                                  double area_tol = 1e-12);
};

/// Builds a matrix from COO triplets (the reference's InterpolationMatrix
/// constructor followed by to_csr()); duplicates are kept.
/// nan_batch_apply (apply.hpp:359-420) on FP64 layout_left fields (device
/// pointers): NaN sources are left out and each row renormalised by its
/// valid weight; NaN where the valid fraction is below 1 - na_thres - 1e-6.
void nan_batch_apply(InterpolationMatrix &A, const double *x, double *y, index_t n_src, index_t n_dst, index_t k, double na_thres,
                     hipStream_t s);
/// Masked apply (apply.hpp:438-484): destination rows with mask 0 keep their
/// values; the mask applies when it has n_dst entries (n_mask == n_dst).
void masked_apply(InterpolationMatrix &A, const double *x, double *y, const std::int32_t *dst_mask, index_t n_mask, index_t n_src,
                  index_t n_dst, index_t k, hipStream_t s);
/// AXWC weight blob (weight_cache.hpp): size, write (buf == nullptr: size
/// query), header validation alone, and read.
std::size_t axwc_size(index_t n_src, index_t n_dst, index_t nnz);
std::size_t serialize_axwc(const InterpolationMatrix &M, void *buf, std::size_t buf_size);
void check_axwc(const void *buf, std::size_t buf_size, index_t &n_src, index_t &n_dst, index_t &nnz);
std::unique_ptr<InterpolationMatrix> deserialize_axwc(const void *buf, std::size_t buf_size);
/// ESMF weight file (esmf_weight_io.cpp), NetCDF classic format: write;
/// read = parse on the host (0-based triplets, indices checked) + build.
void write_esmf(const std::string &path, const InterpolationMatrix &M);
struct EsmfTriplets {
    index_t n_src = 0, n_dst = 0;
    std::vector<index_t> rows, cols;
    std::vector<double> val;
};
EsmfTriplets parse_esmf(const std::string &path);
std::unique_ptr<InterpolationMatrix> matrix_from_esmf(const EsmfTriplets &t);
std::unique_ptr<InterpolationMatrix> matrix_from_coo(const index_t *rows, const index_t *cols, const double *vals, index_t nnz, index_t n_src,
                                                     index_t n_dst, const double *frac_a, const double *frac_b, const double *area_a,
                                                     const double *area_b);

/// Y = A X for device-resident layout_left fields X [n_src, k], Y [n_dst, k].
void batch_apply(InterpolationMatrix &A, const void *x, FieldType xt, void *y, FieldType yt, index_t n_src, index_t n_dst, index_t k,
                 ApplyMode mode, hipStream_t s);

/// batch_apply that allocates nothing: the matrix's apply layout (and, for a
/// staged launch, its tiles) must already exist, and long-row partial sums go
/// to a fixed per-device array of noalloc_scratch_doubles() doubles; larger
/// requests are rejected.  Calls that overlap on one GPU share that array and
/// race.  Enqueues only (no synchronisation, no memset).
void batch_apply_noalloc(InterpolationMatrix &A, const void *x, FieldType xt, void *y, FieldType yt, index_t n_src, index_t n_dst, index_t k,
                         ApplyMode mode, hipStream_t s);
std::size_t noalloc_scratch_doubles();

ConservationReport check_conservation(const InterpolationMatrix &A, const double *d_src, const double *d_dst, NormType norm, hipStream_t s);

/// Non-owning layout_left view of a field array, the shape of the
/// reference's field_view<T, Rank> (types.hpp:69-70: an mdspan with dynamic
/// extents and layout_left): element (i, j) of a rank-2 view is data[i + j *
/// extent(0)], so column j (one field) is contiguous.  Device memory here.
template <class T, std::size_t Rank>
class field_view {
    static_assert(Rank == 1 || Rank == 2, "field_view: rank 1 or 2");

   public:
    field_view() = default;
    field_view(T *data, std::size_t n0) requires(Rank == 1) : p_(data), e_{n0} {}
    field_view(T *data, std::size_t n0, std::size_t n1) requires(Rank == 2) : p_(data), e_{n0, n1} {}
    T *data_handle() const { return p_; }
    std::size_t extent(std::size_t r) const { return r < Rank ? e_[r] : 1; }
    static constexpr std::size_t rank() { return Rank; }

   private:
    T *p_ = nullptr;
    std::size_t e_[Rank] = {};
};

/// apply (apply.hpp:61-119): dst = A src for one FP64 field; extents checked
/// before dst is touched (the reference's messages); returns when done.
void apply(InterpolationMatrix &A, field_view<const double, 1> src, field_view<double, 1> dst, hipStream_t s = nullptr);
/// Masked apply (apply.hpp:457-484): rows with dst_mask 0 keep their values.
void apply(InterpolationMatrix &A, field_view<const double, 1> src, field_view<double, 1> dst, field_view<const std::int32_t, 1> dst_mask,
           hipStream_t s = nullptr);
/// batch_apply (apply.hpp:275-343): dst(:, v) = A src(:, v) for every field v.
void batch_apply(InterpolationMatrix &A, field_view<const double, 2> src, field_view<double, 2> dst, hipStream_t s = nullptr);
/// nan_batch_apply (apply.hpp:359-436) on views.
void nan_batch_apply(InterpolationMatrix &A, field_view<const double, 2> src, field_view<double, 2> dst, double na_thres, hipStream_t s = nullptr);

/// Launch shape of the apply: fields per step, field chunks (more chunks =
/// more workgroups, more re-reads of the matrix) and whether tiles stage
/// their source values in LDS.
struct ApplyShape {
    int fields_per_step = 4, field_chunks = 1;
    bool staged = false;
    bool wave = false;  // staged per wave (wave tiles) rather than per workgroup
};

/// Pins the launch shape batch_apply uses for this (n_dst, nnz, k, types,
/// mode), as the autotuner's cache entry would.
void set_apply_shape(InterpolationMatrix &A, index_t k, FieldType xt, FieldType yt, ApplyMode mode, const ApplyShape &shape);

/// Times the candidate launch shapes of batch_apply on the given device
/// fields and caches the fastest for this (n_dst, nnz, k, types, mode).
ApplyShape autotune_apply(InterpolationMatrix &A, const void *x, FieldType xt, void *y, FieldType yt, index_t k, ApplyMode mode, hipStream_t s);

/// Launch shape of a persistent multi-pair batch: fields per step (staged
/// together in LDS) and fields per work item.
/// A plan's recording: the persistent work-queue kernel with its fields per
/// step and per work item, or (pair_launches) each pair's own autotuned
/// apply launches captured one after another in the graph.
struct BatchShape {
    int fields_per_step = 4, fields_per_item = 16;
    bool pair_launches = false;
};

/// Several regrid pairs applied in one HIP graph: a persistent work-queue
/// launch over (pair, tile, field range) items, or -- when the autotuner
/// measures it faster -- each pair's own tuned apply launches captured in
/// order.  Pair p computes dst[p] = A[p] src[p] on device-resident
/// layout_left fields [n_src, k[p]] -> [n_dst, k[p]]; all pairs share field
/// types and mode.  The field pointers are bound at construction (the graph
/// records them); run() replays the graph on any stream.  Replays of one
/// plan must not overlap.
class BatchPlan {
   public:
    BatchPlan(std::vector<std::shared_ptr<InterpolationMatrix>> mats, std::vector<const void *> src, std::vector<void *> dst,
              std::vector<index_t> k, FieldType xt, FieldType yt, ApplyMode mode, int autotune, hipStream_t s);
    ~BatchPlan();
    void run(hipStream_t s, bool sync);
    BatchShape shape() const;
    index_t n_items() const;
    int n_workgroups() const;

   private:
    struct Impl;
    std::unique_ptr<Impl> impl_;
};

// ---------------------------------------------------------------- internal
namespace detail {

/// Per-mesh data derived inside one generate() call (never cached across calls).
struct MeshGeom {
    index_t n_cells = 0, n_nodes = 0;
    PinnedArray<double> h_xyz;          // node xyz [3 * n_nodes] (libm, reference bits)
    PinnedArray<double> h_xyz_deg;      // radian meshes: xyz of the coordinates read as degrees
    bool xyz_ready = false;             // h_xyz (and h_xyz_deg) filled by node_xyz_host
    bool dev_xyz = false;               // d_xyz built on the device by node_xyz_device (h_xyz unused)
    bool d_xyz_ready = false;           // d_xyz holds h_xyz (uploaded by node_xyz_host; ordered by the caller)
    DeviceBuffer<double> d_xyz;         // same on the device
    DeviceBuffer<double> d_area;        // [n_cells]
    DeviceBuffer<std::int32_t> d_flags; // bit0 degenerate, bit1 masked, bit2 area<=0
    DeviceBuffer<double> d_cap;         // [4 * n_cells] centroid xyz + angular radius
    DeviceBuffer<float> d_box;          // [6 * n_cells] float AABB lo xyz, hi xyz
    // regular-grid node trig cache (trig_cache.hpp:179-267)
    bool regular = false;
    index_t ni = 0, nj = 0;
    double lon_min = 0, delta_lon = 0, lat_min = 0, delta_lat = 0;
    std::vector<double> h_trig;   // [2*(ni+1) + 2*(nj+1)]: sin_lon, cos_lon, sin_lat, cos_lat
    DeviceBuffer<double> d_trig;
};

struct GridAxes {
    bool all_quads = false;
    std::vector<double> ulon, ulat;  // tolerance-unique sorted node coordinates
    std::vector<double> xlon, xlat;  // exact distinct node coordinates, sorted (when the detection collected them)
};

GridAxes unique_axes(const Mesh &m, hipStream_t s);
bool regular_from_axes(const Mesh &m, const GridAxes &ax, MeshGeom &g);
bool rectilinear_from_axes(const Mesh &m, const GridAxes &ax);

/// Node xyz (and, for radian meshes, the degree reading) with host libm;
/// touches only g.h_xyz, g.h_xyz_deg and g.xyz_ready, so it may run on a host
/// thread while the GPU works on other members.
///
/// With `up` given, `d_out` (3 * n_nodes doubles, g.d_xyz's memory taken by
/// the caller before starting this on another thread) receives each host
/// thread's slice of h_xyz on stream `up` as soon as that slice is done, and
/// `done` is recorded on `up` after the last one; returns whether it was
/// (false when cancelled).  The caller orders its stream after `done` and
/// sets g.d_xyz_ready.
bool node_xyz_host(const Mesh &m, MeshGeom &g, const std::atomic<bool> *cancel = nullptr, hipStream_t up = nullptr, hipEvent_t done = nullptr,
                   double *d_out = nullptr);
/// Node xyz of a degree mesh built on the device into d_out [3 * n_nodes] on
/// stream `st`, with host libm's bits: double-double sin/cos give the
/// correctly rounded values, which host libm (within ~0.55 ULP) also returns
/// unless the exact value lies near a rounding midpoint; those angles are
/// evaluated by host libm and patched in.  Waits for `st`.  Returns false
/// (d_out partly written) for a mesh not in degrees, with AXIS_HOST_NODE_XYZ
/// set, with too many coordinates beyond the reduction's range, or when
/// `cancel` is set.
bool node_xyz_trig(const Mesh &m, double *d_out, hipStream_t st, const std::atomic<bool> *cancel = nullptr);
/// A non-blocking stream of the current device for host-to-device uploads
/// beside the caller's stream (created once per device, default priority).
hipStream_t upload_stream();
/// A second such stream: the long rows of the CSR assembly beside the short ones.
hipStream_t rows_stream();
/// xyz of one node as node_xyz_host computes it (same bits).
void node_xyz_one(const Mesh &m, index_t node, double out[3]);
/// Host-only necessary condition for node_xyz_device_eligible: degree
/// coordinates, all cells quads, and node and cell counts of some
/// nlon x nlat grid (nodes = nlon nlat, cells = (nlon - 1)(nlat - 1)).
bool maybe_rectilinear_deg(const Mesh &m);
/// Whether node_xyz_device can serve the mesh: degree coordinates on a
/// rectilinear grid (and AXIS_HOST_NODE_XYZ unset).
bool node_xyz_device_eligible(const Mesh &m, const GridAxes &ax);
/// Node xyz of an eligible mesh built on the device into g.d_xyz from host
/// libm sin/cos tables of the grid's axis values (bit-identical to
/// node_xyz_host); nodes whose coordinates match no axis value bit for bit
/// are computed on the host.  Returns false (and leaves g.dev_xyz unset and
/// g.d_xyz untouched) when the mesh is not eligible or too many nodes need the
/// host.  The caller must not have slice uploads into g.d_xyz still to queue.
/// `tables`: axis_trig_tables(ax), when the caller computed them beforehand.
bool node_xyz_device(const Mesh &m, const GridAxes &ax, MeshGeom &g, hipStream_t s, PinnedArray<double> *tables = nullptr);
/// The host-libm sin/cos tables of a grid's axis values that node_xyz_device
/// uploads (pure host work: may run on another thread meanwhile).
PinnedArray<double> axis_trig_tables(const GridAxes &ax);
void prepare_geometry(const Mesh &m, MeshGeom &g, hipStream_t s);

/// 30-bit Morton code of every cell's centroid (normalised mean of its vertex
/// xyz), used to order the rows of a generated matrix for the apply.
void cell_morton_keys(const Mesh &m, DeviceBuffer<std::uint32_t> &key, hipStream_t s);

// scan / sort primitives (scan.hip, sort.hip)
void exclusive_scan_i64(const index_t *in, index_t *out, index_t n, hipStream_t s);  // out[n] = total
/// Copies `bytes` between device memory and pinned host memory (PinnedCache
/// blocks: mapped, device-accessible) with a copy kernel on `s` instead of a
/// DMA engine.  The node-xyz side stream copies this way: with both the side
/// stream and the generating stream on the DMA engines, a copy of the
/// generating stream was seen to wait ~5 ms about once in 30 generations.
void kernel_copy(void *dst, const void *src, std::size_t bytes, hipStream_t s);
void radix_sort_pairs_u32(std::uint32_t *keys, std::uint32_t *vals, std::size_t n, int bits, hipStream_t s);
/// Stable LSD sort on key bits [first_bit, bits) (first_bit a multiple of 8):
/// keys whose lower bits are already ascending need only the upper passes.
void radix_sort_keys_u64(std::uint64_t *keys, std::size_t n, hipStream_t s, int bits = 64, int first_bit = 0);
void radix_sort_pairs_u64(std::uint64_t *keys, std::uint32_t *vals, std::size_t n, int bits, hipStream_t s);

// LBVH broad phase (lbvh.hip): candidate CSR of src cells whose float boxes
// intersect each dst cell's box.
/// Source-side LBVH (packed internal nodes); the source boxes must outlive it.
struct LbvhTree {
    int n = 0;
    const float *src_boxes = nullptr;
    DeviceBuffer<std::uint8_t> nodes;
};
/// Broad phase of a rectilinear degree source through its (lon, lat) cell
/// grid (gridquery.hip): the same candidate sets as lbvh_query.
struct GridIndex {
    bool ready = false;
    bool grid_order = false;               // cell j (nlon - 1) + i is slot (i, j) (cell_at is the identity)
    int nlon = 0, nlat = 0;                // node counts per axis
    double ulon0 = 0.0, dlon = 0.0;        // first longitude, mean spacing
    DeviceBuffer<double> axes;             // tolerance-unique lon values, then lat values
    DeviceBuffer<std::int32_t> cell_at;    // source cell of grid slot j (nlon - 1) + i
    DeviceBuffer<float> rows;              // per cell row: z-band suffix minima, prefix maxima, longitude overhang
    const float *src_boxes = nullptr;
};
/// Builds the grid of a source whose axes passed rectilinear_from_axes;
/// false (index not ready) when its cells are not the 2 x 2 node blocks of
/// that grid, each once, or its tables exceed 64 KiB of LDS.  Waits for `s`.
bool grid_index_build(const Mesh &src, const GridAxes &ax, const float *src_boxes, GridIndex &G, hipStream_t s);
/// Candidates (cand_off [n_dst + 1], cand_src) and each one's destination row
/// (pair_row); returns their number.  Waits for `s`.
index_t grid_query(const GridIndex &G, const float *dst_boxes, index_t n_dst, DeviceBuffer<index_t> &cand_off, DeviceBuffer<std::int32_t> &cand_src,
                   DeviceBuffer<std::int32_t> &pair_row, hipStream_t s);
void lbvh_build(const float *src_boxes, index_t n_src, LbvhTree &t, hipStream_t s);
void lbvh_query(const LbvhTree &t, const float *dst_boxes, index_t n_dst, DeviceBuffer<index_t> &cand_off, DeviceBuffer<std::int32_t> &cand_src,
                hipStream_t s);
void lbvh_candidates(const float *src_boxes, index_t n_src, const float *dst_boxes, index_t n_dst, DeviceBuffer<index_t> &cand_off,
                     DeviceBuffer<std::int32_t> &cand_src, hipStream_t s);

/// WeightGenerator::generate's longitude-disjoint check in two halves:
/// launch() queues both meshes' longitude ranges and their readback on `s`
/// without waiting; finish() waits for that readback (normally already done
/// by a wait of the generation itself) and throws the reference's mismatch
/// error.  A generation that meets an error of its own before finish() runs
/// reports that one instead.
class LonCheck {
   public:
    void launch(const Mesh &src, const Mesh &dst, hipStream_t s);
    void finish(bool structured_fast_path);
    ~LonCheck();

   private:
    const Mesh *src_ = nullptr, *dst_ = nullptr;
    DeviceBuffer<unsigned long long> d_;
    PinnedArray<unsigned long long> h_;
    hipEvent_t done_ = nullptr;
    bool armed_ = false;
};

/// First-order conservative weights; with a proposer network the broad
/// phase is the network's guarded search caps, with LBVH fallback rows.
/// `lon` (launched by the caller) is finished after the source's grid
/// detection.
std::unique_ptr<InterpolationMatrix> generate_conservative(const Mesh &src, const Mesh &dst, const RegridConfig &cfg, hipStream_t s,
                                                           const mlp::Model *model = nullptr, ProposerStats *pst = nullptr,
                                                           double area_tol = 1e-12, LonCheck *lon = nullptr);
std::unique_ptr<InterpolationMatrix> generate_bilinear_rect(const Mesh &src, const Mesh &dst, const GridAxes &src_axes, hipStream_t s);
void check_longitude_ranges(const Mesh &src, const Mesh &dst, bool structured_fast_path, hipStream_t s);
/// coastal_renormalize_and_extrapolate (weight_generator.cpp:1736-1881):
/// with a source mask, keeps the wet weights (renormalised where their row
/// sum is in (0, 1)) and, for NearestWet, maps each row left empty to its
/// nearest wet source cell with weight 1.  No-op without a source mask.
void coastal_pass(const Mesh &src, const Mesh &dst, ExtrapolationAction extrap, InterpolationMatrix &M, hipStream_t s);
/// Analytic paths for structured pairs (null when the path declines):
/// generate_conservative_rect (weight_generator_conservative_rect.cpp:87-395),
/// generate_conservative_rect_nonuniform
/// (weight_generator_conservative_rect_nonuniform.cpp:52-313) and
/// generate_bilinear_rect (weight_generator_bilinear_rect.cpp:182-266).
std::unique_ptr<InterpolationMatrix> conservative_regular(const Mesh &src, const Mesh &dst, const RegridConfig &cfg, const GridAxes &sax,
                                                          const GridAxes &dax, const MeshGeom &sg, const MeshGeom &dg, hipStream_t s);
std::unique_ptr<InterpolationMatrix> conservative_rectilinear(const Mesh &src, const Mesh &dst, const RegridConfig &cfg, const GridAxes &sax,
                                                              const GridAxes &dax, hipStream_t s);
std::unique_ptr<InterpolationMatrix> bilinear_regular(const Mesh &src, const Mesh &dst, const RegridConfig &cfg, const GridAxes &sax, const MeshGeom &sg,
                                                      hipStream_t s);
/// NearestNeighbor weights (generate_nearest, weight_generator.cpp:2025-2271):
/// the analytic index rounding when both meshes are regular lat-lon grids
/// (src_reg and dst_reg given), else the nearest source centroid point.
std::unique_ptr<InterpolationMatrix> generate_nearest(const Mesh &src, const Mesh &dst, const RegridConfig &cfg, const MeshGeom *src_reg,
                                                      const MeshGeom *dst_reg, hipStream_t s);
/// Distributed generate of one rank (weight_generator.cpp:4525-4680): the
/// local matrix with its off-rank columns remapped to n_local + slot, and
/// the HaloPattern of those columns.  owner_of_src maps a global source id
/// to its rank; this rank is the owner of the first local source cell.
std::unique_ptr<InterpolationMatrix> generate_distributed(const Mesh &src, const Mesh &dst, const RegridConfig &cfg, const index_t *src_global_ids,
                                                          index_t n_src_ids, const std::int32_t *owner_of_src, index_t n_owner, HaloPattern &pattern,
                                                          hipStream_t s);
/// The gathered-buffer distributed apply (apply.hpp:151-211) on FP64
/// layout_left device fields: dst = S [local_src ; halo].
void apply_distributed(InterpolationMatrix &A, const HaloPattern &pattern, const double *local_src, index_t n_local, const double *halo,
                       index_t n_halo, double *dst, index_t n_dst, index_t k, hipStream_t s);
/// Float (lon, lat) centroids of every cell (compute_cell_centroids_xy,
/// weight_generator.cpp:74-138, rounded to float) into device memory, equal
/// to host libm's; returns how many cells were recomputed on the host.
index_t cell_centroids_f32(const Mesh &m, float2 *out, hipStream_t s);

}  // namespace detail
}  // namespace axis_hip
