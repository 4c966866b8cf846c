// bilinear.hip -- bilinear weights from a rectilinear lat-lon source, and the
// WeightGenerator::generate dispatch.
//
// generate_bilinear_rect_nonuniform (libs/axis/src/solver/
// weight_generator_bilinear_rect_nonuniform.cpp:22-149) as one thread per
// destination cell: destination point = arithmetic mean of the vertex lon/lat
// degrees (no periodic wrap), upper_bound in the unique cell-centre arrays,
// clamped stencil, clamped tx/ty, exactly four entries per row (zeros kept),
// frac = 1, areas 0.  Only + - * / and comparisons: bit-exact.
#include <hip/hip_runtime.h>

#include <chrono>

#include "axis_hip.hpp"

namespace axis_hip {
namespace detail {

namespace {

__device__ inline int upper_bound_idx(const double *a, int n, double x) {
    int lo = 0, hi = n;  // first index with a[idx] > x
    while (lo < hi) {
        int mid = lo + ((hi - lo) >> 1);
        if (x < a[mid])
            hi = mid;
        else
            lo = mid + 1;
    }
    return lo;
}

__global__ void bilinear_kernel(const index_t *off, const index_t *ind, const double *coords, index_t n_nodes, index_t n_dst,
                                const double *clon, int ni, const double *clat, int nj, index_t *row_ptr, index_t *col, double *val) {
    index_t c = static_cast<index_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (c > n_dst) return;
    row_ptr[c] = 4 * c;
    if (c == n_dst) return;
    const index_t b = off[c], e = off[c + 1];
    double ls = 0.0, as = 0.0;
    for (index_t k = b; k < e; ++k) {
        index_t nd = ind[k];
        ls += coords[nd];
        as += coords[nd + n_nodes];
    }
    const double lon = ls / static_cast<double>(e - b);
    const double lat = as / static_cast<double>(e - b);
    int i = upper_bound_idx(clon, ni, lon) - 1;
    int j = upper_bound_idx(clat, nj, lat) - 1;
    if (i < 0) i = 0;
    if (i >= ni - 1) i = ni - 2;
    if (i < 0) i = 0;
    if (j < 0) j = 0;
    if (j >= nj - 1) j = nj - 2;
    if (j < 0) j = 0;
    const double x0 = clon[i], x1 = clon[i + 1], y0 = clat[j], y1 = clat[j + 1];
    double tx = (lon - x0) / (x1 - x0);
    double ty = (lat - y0) / (y1 - y0);
    tx = fmax(0.0, fmin(1.0, tx));
    ty = fmax(0.0, fmin(1.0, ty));
    const int i1 = i + 1, j1 = j + 1;
    index_t idx[4] = {static_cast<index_t>(j) * ni + i, static_cast<index_t>(j) * ni + i1, static_cast<index_t>(j1) * ni + i,
                      static_cast<index_t>(j1) * ni + i1};
    double w[4] = {(1.0 - tx) * (1.0 - ty), tx * (1.0 - ty), (1.0 - tx) * ty, tx * ty};
    // to_csr order: ascending column (already so unless ni < 2)
    for (int a = 1; a < 4; ++a)
        for (int q = a; q > 0 && idx[q - 1] > idx[q]; --q) {
            index_t ti = idx[q];
            idx[q] = idx[q - 1];
            idx[q - 1] = ti;
            double tw = w[q];
            w[q] = w[q - 1];
            w[q - 1] = tw;
        }
    for (int a = 0; a < 4; ++a) {
        col[4 * c + a] = idx[a];
        val[4 * c + a] = w[a];
    }
}

__global__ void fill_kernel(double *p, index_t n, double v) {
    index_t i = static_cast<index_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

}  // namespace

std::unique_ptr<InterpolationMatrix> generate_bilinear_rect(const Mesh &src, const Mesh &dst, const GridAxes &ax, hipStream_t s) {
    const auto t0 = std::chrono::steady_clock::now();
    const index_t n_src = src.n_cells(), n_dst = dst.n_cells();
    const int ni = static_cast<int>(ax.ulon.size()) - 1, nj = static_cast<int>(ax.ulat.size()) - 1;
    // the reference reads centre i + 1 unconditionally, which needs two cells per axis
    if (ni < 2 || nj < 2) throw InvalidArgument("generate_bilinear: the rectilinear source needs at least 2 cells along each axis");
    std::vector<double> clon(ni), clat(nj);
    for (int i = 0; i < ni; ++i) clon[i] = 0.5 * (ax.ulon[i] + ax.ulon[i + 1]);
    for (int j = 0; j < nj; ++j) clat[j] = 0.5 * (ax.ulat[j] + ax.ulat[j + 1]);
    if (n_dst > 0 && dst.min_cell_vertices() == 0) throw InvalidArgument("generate_bilinear: destination cell without vertices");
    DeviceBuffer<double> d_clon(std::max(ni, 1)), d_clat(std::max(nj, 1));
    d_clon.upload(clon.data(), ni, s);
    d_clat.upload(clat.data(), nj, s);
    auto M = std::make_unique<InterpolationMatrix>();
    M->n_src = n_src;
    M->n_dst = n_dst;
    M->nnz = 4 * n_dst;
    M->row_ptr.resize(n_dst + 1);
    M->col.resize(std::max<index_t>(4 * n_dst, 1));
    M->val.resize(std::max<index_t>(4 * n_dst, 1));
    M->frac_a.resize(std::max<index_t>(n_src, 1));
    M->frac_b.resize(std::max<index_t>(n_dst, 1));
    M->area_a.resize(std::max<index_t>(n_src, 1));
    M->area_b.resize(std::max<index_t>(n_dst, 1));
    M->max_row_nnz = n_dst > 0 ? 4 : 0;
    M->len_hist[4] = n_dst;  // every row holds its four stencil entries
    M->len_hist_ready = true;
    cell_morton_keys(dst, M->row_key, s);
    bilinear_kernel<<<grid_for(n_dst + 1, 256), 256, 0, s>>>(dst.d_offsets.get(), dst.d_indices.get(), dst.d_coords.get(), dst.n_nodes(), n_dst,
                                                             d_clon.get(), ni, d_clat.get(), nj, M->row_ptr.get(), M->col.get(), M->val.get());
    AXIS_LAUNCH_CHECK();
    if (n_src > 0) fill_kernel<<<grid_for(n_src, 256), 256, 0, s>>>(M->frac_a.get(), n_src, 1.0);
    if (n_dst > 0) fill_kernel<<<grid_for(n_dst, 256), 256, 0, s>>>(M->frac_b.get(), n_dst, 1.0);
    AXIS_HIP_CHECK(hipMemsetAsync(M->area_a.get(), 0, sizeof(double) * std::max<index_t>(n_src, 1), s));
    AXIS_HIP_CHECK(hipMemsetAsync(M->area_b.get(), 0, sizeof(double) * std::max<index_t>(n_dst, 1), s));
    AXIS_HIP_CHECK(hipStreamSynchronize(s));
    M->gen_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return M;
}

}  // namespace detail

// WeightGenerator::generate, weight_generator.cpp:1887-2019: coordinate-system
// check, longitude-disjoint check, method dispatch, then the masked-source
// post-pass (:2018), which returns its input when the source has no mask.
std::unique_ptr<InterpolationMatrix> WeightGenerator::generate(const Mesh &src, const Mesh &dst, const RegridConfig &cfg, hipStream_t s) {
    return generate_proposed(src, dst, cfg, nullptr, nullptr, s);
}

std::unique_ptr<InterpolationMatrix> WeightGenerator::generate_proposed(const Mesh &src, const Mesh &dst, const RegridConfig &cfg,
                                                                        const mlp::Model *proposer, ProposerStats *stats, hipStream_t s,
                                                                        double area_tol) {
    if (proposer && cfg.method != InterpolationMethod::Conservative1stOrder)
        throw InvalidArgument("WeightGenerator::generate: the learned proposer serves Conservative1stOrder only");
    if (src.coord_system() != dst.coord_system())
        throw InvalidArgument("WeightGenerator::generate: source and destination meshes have different CoordinateSystem values (src=" +
                              std::to_string(static_cast<int>(src.coord_system())) + ", dst=" + std::to_string(static_cast<int>(dst.coord_system())) +
                              ")");
    const bool cartesian = src.coord_system() == CoordinateSystem::Cartesian3D;
    std::unique_ptr<InterpolationMatrix> M;
    switch (cfg.method) {
        case InterpolationMethod::Conservative1stOrder: {
            detail::LonCheck lon;  // (finished after the source's grid detection)
            lon.launch(src, dst, s);
            M = detail::generate_conservative(src, dst, cfg, s, proposer, stats, area_tol, &lon);
            lon.finish(false);  // (a no-op once finished)
            break;
        }
        case InterpolationMethod::Bilinear: {
            detail::LonCheck lon;
            lon.launch(src, dst, s);
            detail::GridAxes sax = detail::unique_axes(src, s);
            const bool s_rect = detail::rectilinear_from_axes(src, sax);
            lon.finish(s_rect);
            detail::MeshGeom sg, dg;
            if (detail::regular_from_axes(src, sax, sg)) {
                detail::GridAxes dax = detail::unique_axes(dst, s);
                if (detail::regular_from_axes(dst, dax, dg)) M = detail::bilinear_regular(src, dst, cfg, sax, sg, s);
            }
            if (M) break;
            if (cartesian)
                throw InvalidArgument("WeightGenerator::generate: a Cartesian3D bilinear pair is served by the regular-grid path only here");
            if (!s_rect)
                throw InvalidArgument("WeightGenerator::generate: bilinear needs a rectilinear lat-lon source here (the reference's BVH "
                                      "bilinear path is not provided)");
            M = detail::generate_bilinear_rect(src, dst, sax, s);
            break;
        }
        case InterpolationMethod::NearestNeighbor: {
            // generate_nearest (host dispatch, weight_generator.cpp:2107-2127)
            detail::check_longitude_ranges(src, dst, false, s);
            detail::GridAxes sax = detail::unique_axes(src, s);
            detail::MeshGeom sg, dg;
            bool reg = detail::regular_from_axes(src, sax, sg);
            if (reg) {
                detail::GridAxes dax = detail::unique_axes(dst, s);
                reg = detail::regular_from_axes(dst, dax, dg);
            }
            M = detail::generate_nearest(src, dst, cfg, reg ? &sg : nullptr, reg ? &dg : nullptr, s);
            break;
        }
        default:
            throw InvalidArgument("WeightGenerator::generate: only Bilinear, NearestNeighbor and Conservative1stOrder are provided");
    }
    const auto t0 = std::chrono::steady_clock::now();
    detail::coastal_pass(src, dst, cfg.extrap_method, *M, s);  // :2018
    M->gen_ms += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return M;
}

}  // namespace axis_hip
