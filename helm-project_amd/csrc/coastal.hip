// Masked-source post-pass of WeightGenerator::generate
// (coastal_renormalize_and_extrapolate, weight_generator.cpp:1736-1881).
//
// When the source mesh carries a cell mask:
//   * only weights whose source cell is wet (mask > 0) are kept, and a row's
//     kept weights are divided by their sum when that sum lies in (0, 1);
//   * with ExtrapolationAction::NearestWet, a row left without weights gets
//     one weight 1.0 on the wet (mask != 0) source cell whose (lon, lat)
//     centroid, rounded to float, is nearest in the plane (ArborX nearest on
//     Point<2>, :1801-1855).
// Fractions, areas and the rows' spatial keys are unchanged (:1879).
//
// GPU steps, all on the generation stream:
//   1. wet_rows_kernel      per row: wet count and wet-weight sum (CSR order)
//   2. host                 dry rows; float centroids of wet sources and dry
//                           rows with host libm (the reference computes them
//                           on the host with the same formulas)
//   3. point tree           2D Morton sort of the wet points, leaves of 32
//                           points, implicit complete binary tree of float
//                           boxes (heap order, root 1)
//   4. nearest_kernel       one thread per dry row: exact nearest (smallest
//                           float squared distance, then smallest source
//                           index), pruning boxes whose float distance is
//                           larger than the best
//   5. new row counts -> exclusive scan -> coastal_fill_kernel
#include <algorithm>
#include <cmath>
#include <limits>

#include "axis_hip.hpp"
#include "host_util.hpp"

namespace axis_hip::detail {
namespace {

constexpr int kLeafPts = 32;
constexpr int kStack = 64;

__global__ void wet_rows_kernel(const index_t *__restrict__ row_ptr, const index_t *__restrict__ col, const double *__restrict__ val,
                                const std::int32_t *__restrict__ mask, index_t n_rows, std::int32_t *__restrict__ wet_cnt,
                                double *__restrict__ wet_sum) {
    const index_t r = static_cast<index_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (r >= n_rows) return;
    double s = 0.0;
    std::int32_t c = 0;
    for (index_t e = row_ptr[r]; e < row_ptr[r + 1]; ++e) {
        if (mask[col[e]] > 0) {
            s += val[e];
            ++c;
        }
    }
    wet_cnt[r] = c;
    wet_sum[r] = s;
}

// 32-bit 2D Morton code of 16-bit quantised coordinates.
__device__ inline std::uint32_t spread16(std::uint32_t v) {
    v &= 0xFFFFu;
    v = (v | (v << 8)) & 0x00FF00FFu;
    v = (v | (v << 4)) & 0x0F0F0F0Fu;
    v = (v | (v << 2)) & 0x33333333u;
    v = (v | (v << 1)) & 0x55555555u;
    return v;
}

__global__ void point_keys_kernel(const float *__restrict__ px, const float *__restrict__ py, index_t n, float x0, float y0, float sx, float sy,
                                  std::uint32_t *__restrict__ key, std::uint32_t *__restrict__ pos) {
    const index_t q = static_cast<index_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (q >= n) return;
    const float fx = fminf(fmaxf((px[q] - x0) * sx, 0.0f), 65535.0f);
    const float fy = fminf(fmaxf((py[q] - y0) * sy, 0.0f), 65535.0f);
    key[q] = spread16(static_cast<std::uint32_t>(fx)) | (spread16(static_cast<std::uint32_t>(fy)) << 1);
    pos[q] = static_cast<std::uint32_t>(q);
}

__global__ void gather_points_kernel(const std::uint32_t *__restrict__ pos, const float *__restrict__ px, const float *__restrict__ py,
                                     const std::int32_t *__restrict__ pidx, index_t n, float2 *__restrict__ pts, std::int32_t *__restrict__ idx) {
    const index_t q = static_cast<index_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (q >= n) return;
    const std::uint32_t p = pos[q];
    pts[q] = make_float2(px[p], py[p]);
    idx[q] = pidx[p];
}

// Node boxes as (xmin, ymin, xmax, ymax); a leaf past the last point gets an
// inverted box (never entered: the traversal checks the leaf range).
__global__ void leaf_boxes_kernel(const float2 *__restrict__ pts, index_t n, index_t n_leaves, index_t P, float4 *__restrict__ box) {
    const index_t l = static_cast<index_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (l >= P) return;
    float4 b = make_float4(INFINITY, INFINITY, -INFINITY, -INFINITY);
    if (l < n_leaves) {
        const index_t e = std::min<index_t>(n, (l + 1) * kLeafPts);
        for (index_t q = l * kLeafPts; q < e; ++q) {
            const float2 p = pts[q];
            b.x = fminf(b.x, p.x);
            b.y = fminf(b.y, p.y);
            b.z = fmaxf(b.z, p.x);
            b.w = fmaxf(b.w, p.y);
        }
    }
    box[P + l] = b;
}

__global__ void inner_boxes_kernel(float4 *__restrict__ box, index_t lo) {
    const index_t i = lo + static_cast<index_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= 2 * lo) return;
    const float4 a = box[2 * i], b = box[2 * i + 1];
    box[i] = make_float4(fminf(a.x, b.x), fminf(a.y, b.y), fmaxf(a.z, b.z), fmaxf(a.w, b.w));
}

// Float squared distance from q to a box.  For any point p inside the box the
// rounded differences and squares are monotone in the exact ones, so this is
// never above the float squared distance to p: pruning on "box > best" keeps
// every point that could win or tie.
__device__ inline float box_d2(const float4 b, float qx, float qy) {
    const float dx = fmaxf(fmaxf(b.x - qx, qx - b.z), 0.0f);
    const float dy = fmaxf(fmaxf(b.y - qy, qy - b.w), 0.0f);
    return dx * dx + dy * dy;
}

__global__ void nearest_kernel(const float4 *__restrict__ box, const float2 *__restrict__ pts, const std::int32_t *__restrict__ idx, index_t n,
                               index_t n_leaves, index_t P, const float2 *__restrict__ query, const std::int32_t *__restrict__ qrow, index_t n_q,
                               std::int32_t *__restrict__ nearest) {
    const index_t t = static_cast<index_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= n_q) return;
    const float qx = query[t].x, qy = query[t].y;
    float best = INFINITY;
    std::int32_t bi = std::numeric_limits<std::int32_t>::max();
    index_t stack[kStack];
    int sp = 0;
    stack[sp++] = 1;
    while (sp > 0) {
        const index_t node = stack[--sp];
        if (node >= P) {
            const index_t l = node - P;
            if (l >= n_leaves) continue;
            const index_t e = std::min<index_t>(n, (l + 1) * kLeafPts);
            for (index_t q = l * kLeafPts; q < e; ++q) {
                const float2 p = pts[q];
                const float dx = p.x - qx, dy = p.y - qy;
                const float d2 = dx * dx + dy * dy;
                const std::int32_t id = idx[q];
                if (d2 < best || (d2 == best && id < bi)) {
                    best = d2;
                    bi = id;
                }
            }
            continue;
        }
        const index_t c0 = 2 * node, c1 = 2 * node + 1;
        const float d0 = box_d2(box[c0], qx, qy), d1 = box_d2(box[c1], qx, qy);
        // nearer child on top of the stack
        const index_t near_c = d0 <= d1 ? c0 : c1, far_c = d0 <= d1 ? c1 : c0;
        const float near_d = fminf(d0, d1), far_d = fmaxf(d0, d1);
        if (!(far_d > best) && sp < kStack) stack[sp++] = far_c;
        if (!(near_d > best) && sp < kStack) stack[sp++] = near_c;
    }
    // every point's distance NaN (a NaN centroid): no extrapolated weight
    if (bi != std::numeric_limits<std::int32_t>::max()) nearest[qrow[t]] = bi;
}

__global__ void fill_int_kernel(std::int32_t *__restrict__ p, index_t n, std::int32_t v) {
    const index_t i = static_cast<index_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

__global__ void new_counts_kernel(const std::int32_t *__restrict__ wet_cnt, const std::int32_t *__restrict__ nearest, index_t n_rows,
                                  index_t *__restrict__ cnt) {
    const index_t r = static_cast<index_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (r >= n_rows) return;
    cnt[r] = wet_cnt[r] > 0 ? wet_cnt[r] : (nearest[r] >= 0 ? 1 : 0);
}

__global__ void coastal_fill_kernel(const index_t *__restrict__ row_ptr, const index_t *__restrict__ col, const double *__restrict__ val,
                                    const std::int32_t *__restrict__ mask, const std::int32_t *__restrict__ wet_cnt,
                                    const double *__restrict__ wet_sum, const std::int32_t *__restrict__ nearest, index_t n_rows,
                                    const index_t *__restrict__ new_ptr, index_t *__restrict__ new_col, double *__restrict__ new_val) {
    const index_t r = static_cast<index_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (r >= n_rows) return;
    index_t o = new_ptr[r];
    if (wet_cnt[r] > 0) {
        const double s = wet_sum[r];
        const bool renorm = s > 0.0 && s < 1.0;
        for (index_t e = row_ptr[r]; e < row_ptr[r + 1]; ++e) {
            const index_t c = col[e];
            if (mask[c] > 0) {
                double w = val[e];
                if (renorm) w /= s;
                new_col[o] = c;
                new_val[o] = w;
                ++o;
            }
        }
    } else if (nearest[r] >= 0) {
        new_col[o] = nearest[r];
        new_val[o] = 1.0;
    }
}

// compute_cell_centroids_xy, weight_generator.cpp:74-138: on a sphere the
// mean of the vertices' unit vectors (degrees converted as lon * pi / 180),
// back to lon in [0, 2 pi) and lat in the mesh's unit; otherwise the mean of
// the first two coordinates.
void centroid_lonlat(const Mesh &m, index_t c, SinCosMemo &memo, double &cx, double &cy) {
    const double pi = 3.14159265358979323846;
    const index_t n_nodes = m.n_nodes();
    const double *lon = m.h_coords.data(), *lat = m.h_coords.data() + n_nodes;
    const index_t b = m.h_offsets[c], e = m.h_offsets[c + 1];
    const double inv = (e > b) ? 1.0 / static_cast<double>(e - b) : 0.0;
    if (m.coord_system() == CoordinateSystem::Cartesian3D) {
        double sx = 0.0, sy = 0.0;
        for (index_t k = b; k < e; ++k) {
            sx += lon[m.h_indices[k]];
            sy += lat[m.h_indices[k]];
        }
        cx = sx * inv;
        cy = sy * inv;
        return;
    }
    const bool deg = m.coord_system() == CoordinateSystem::SphericalDeg;
    double sx = 0.0, sy = 0.0, sz = 0.0;
    for (index_t k = b; k < e; ++k) {
        double lo = lon[m.h_indices[k]], la = lat[m.h_indices[k]];
        if (deg) {
            lo = lo * pi / 180.0;
            la = la * pi / 180.0;
        }
        const geo::V3 p = memo.xyz(lo, la);
        sx += p.x;
        sy += p.y;
        sz += p.z;
    }
    double lat_avg = std::asin(sz * inv);
    double lon_avg = std::atan2(sy * inv, sx * inv);
    if (lon_avg < 0.0) lon_avg += 2.0 * pi;
    if (deg) {
        lon_avg = lon_avg * 180.0 / pi;
        lat_avg = lat_avg * 180.0 / pi;
    }
    cx = lon_avg;
    cy = lat_avg;
}

}  // namespace

void coastal_pass(const Mesh &src, const Mesh &dst, ExtrapolationAction extrap, InterpolationMatrix &M, hipStream_t s) {
    if (!src.has_mask()) return;
    const index_t n_dst = M.n_dst, n_src = M.n_src;
    if (n_dst == 0) return;
    PhaseTrace tr("coastal", s);
    const index_t nr = std::max<index_t>(n_dst, 1);
    DeviceBuffer<std::int32_t> wet_cnt(nr), nearest(nr);
    DeviceBuffer<double> wet_sum(nr);
    wet_rows_kernel<<<grid_for(n_dst, 256), 256, 0, s>>>(M.row_ptr.get(), M.col.get(), M.val.get(), src.d_mask.get(), n_dst, wet_cnt.get(),
                                                         wet_sum.get());
    AXIS_LAUNCH_CHECK();
    fill_int_kernel<<<grid_for(n_dst, 256), 256, 0, s>>>(nearest.get(), n_dst, -1);
    AXIS_LAUNCH_CHECK();
    std::vector<std::int32_t> h_cnt(n_dst);
    wet_cnt.download(h_cnt.data(), n_dst, s);

    if (extrap == ExtrapolationAction::NearestWet) {
        // wet sources (mask != 0) in index order, with their float centroids
        std::vector<std::int32_t> wet;
        wet.reserve(n_src);
        for (index_t i = 0; i < n_src; ++i)
            if (src.h_mask[i] != 0) wet.push_back(static_cast<std::int32_t>(i));
        const index_t n_wet = static_cast<index_t>(wet.size());
        if (n_wet > 0) {
            PinnedArray<float> h_px, h_py;
            h_px.resize(n_wet);
            h_py.resize(n_wet);
            parallel_range(n_wet, [&](index_t b, index_t e) {
                SinCosMemo memo;
                for (index_t q = b; q < e; ++q) {
                    double cx, cy;
                    centroid_lonlat(src, wet[q], memo, cx, cy);
                    h_px.data()[q] = static_cast<float>(cx);
                    h_py.data()[q] = static_cast<float>(cy);
                }
            });
            AXIS_HIP_CHECK(hipStreamSynchronize(s));  // h_cnt
            tr.mark("wet_rows+src_centroids");
            std::vector<std::int32_t> dry;
            for (index_t r = 0; r < n_dst; ++r)
                if (h_cnt[r] == 0) dry.push_back(static_cast<std::int32_t>(r));
            const index_t n_q = static_cast<index_t>(dry.size());
            if (n_q > 0) {
                PinnedArray<float2> h_q;
                h_q.resize(n_q);
                parallel_range(n_q, [&](index_t b, index_t e) {
                    SinCosMemo memo;
                    for (index_t t = b; t < e; ++t) {
                        double cx, cy;
                        centroid_lonlat(dst, dry[t], memo, cx, cy);
                        h_q.data()[t] = make_float2(static_cast<float>(cx), static_cast<float>(cy));
                    }
                });
                float x0 = INFINITY, y0 = INFINITY, x1 = -INFINITY, y1 = -INFINITY;
                for (index_t q = 0; q < n_wet; ++q) {
                    x0 = std::min(x0, h_px.data()[q]);
                    x1 = std::max(x1, h_px.data()[q]);
                    y0 = std::min(y0, h_py.data()[q]);
                    y1 = std::max(y1, h_py.data()[q]);
                }
                tr.mark("dry_centroids");
                DeviceBuffer<float> d_px(n_wet), d_py(n_wet);
                DeviceBuffer<std::int32_t> d_wet(n_wet), d_idx(n_wet), d_qrow(n_q);
                DeviceBuffer<float2> d_pts(n_wet), d_q(n_q);
                DeviceBuffer<std::uint32_t> d_key(n_wet), d_pos(n_wet);
                d_px.upload(h_px.data(), n_wet, s);
                d_py.upload(h_py.data(), n_wet, s);
                d_wet.upload(wet.data(), n_wet, s);
                d_q.upload(h_q.data(), n_q, s);
                d_qrow.upload(dry.data(), n_q, s);
                // Morton order only groups the points; the boxes come from
                // the points themselves, so any quantisation is correct.
                const float sx = (x1 > x0 && std::isfinite(x1 - x0)) ? 65535.0f / (x1 - x0) : 0.0f;
                const float sy = (y1 > y0 && std::isfinite(y1 - y0)) ? 65535.0f / (y1 - y0) : 0.0f;
                point_keys_kernel<<<grid_for(n_wet, 256), 256, 0, s>>>(d_px.get(), d_py.get(), n_wet, std::isfinite(x0) ? x0 : 0.0f,
                                                                       std::isfinite(y0) ? y0 : 0.0f, sx, sy, d_key.get(), d_pos.get());
                AXIS_LAUNCH_CHECK();
                radix_sort_pairs_u32(d_key.get(), d_pos.get(), static_cast<std::size_t>(n_wet), 32, s);
                gather_points_kernel<<<grid_for(n_wet, 256), 256, 0, s>>>(d_pos.get(), d_px.get(), d_py.get(), d_wet.get(), n_wet, d_pts.get(),
                                                                          d_idx.get());
                AXIS_LAUNCH_CHECK();
                const index_t n_leaves = (n_wet + kLeafPts - 1) / kLeafPts;
                index_t P = 1;
                while (P < n_leaves) P *= 2;
                DeviceBuffer<float4> d_box(2 * P);
                leaf_boxes_kernel<<<grid_for(P, 256), 256, 0, s>>>(d_pts.get(), n_wet, n_leaves, P, d_box.get());
                AXIS_LAUNCH_CHECK();
                for (index_t lo = P / 2; lo >= 1; lo /= 2) {
                    inner_boxes_kernel<<<grid_for(lo, 256), 256, 0, s>>>(d_box.get(), lo);
                    AXIS_LAUNCH_CHECK();
                }
                nearest_kernel<<<grid_for(n_q, 128), 128, 0, s>>>(d_box.get(), d_pts.get(), d_idx.get(), n_wet, n_leaves, P, d_q.get(),
                                                                  d_qrow.get(), n_q, nearest.get());
                AXIS_LAUNCH_CHECK();
                AXIS_HIP_CHECK(hipStreamSynchronize(s));  // host arrays above are released on return
                tr.mark("nearest");
            }
        }
    }
    AXIS_HIP_CHECK(hipStreamSynchronize(s));
    DeviceBuffer<index_t> cnt(nr + 1), new_ptr(nr + 1);
    new_counts_kernel<<<grid_for(n_dst, 256), 256, 0, s>>>(wet_cnt.get(), nearest.get(), n_dst, cnt.get());
    AXIS_LAUNCH_CHECK();
    exclusive_scan_i64(cnt.get(), new_ptr.get(), n_dst, s);
    index_t new_nnz = 0;
    AXIS_HIP_CHECK(hipMemcpyAsync(&new_nnz, new_ptr.get() + n_dst, sizeof(index_t), hipMemcpyDeviceToHost, s));
    AXIS_HIP_CHECK(hipStreamSynchronize(s));
    DeviceBuffer<index_t> new_col(std::max<index_t>(new_nnz, 1));
    DeviceBuffer<double> new_val(std::max<index_t>(new_nnz, 1));
    coastal_fill_kernel<<<grid_for(n_dst, 256), 256, 0, s>>>(M.row_ptr.get(), M.col.get(), M.val.get(), src.d_mask.get(), wet_cnt.get(),
                                                             wet_sum.get(), nearest.get(), n_dst, new_ptr.get(), new_col.get(), new_val.get());
    AXIS_LAUNCH_CHECK();
    int max_row = 0;
    for (index_t r = 0; r < n_dst; ++r) max_row = std::max(max_row, h_cnt[r] > 0 ? static_cast<int>(h_cnt[r]) : 1);
    AXIS_HIP_CHECK(hipStreamSynchronize(s));
    tr.mark("assemble");
    M.row_ptr = std::move(new_ptr);
    M.col = std::move(new_col);
    M.val = std::move(new_val);
    M.nnz = new_nnz;
    M.max_row_nnz = new_nnz > 0 ? max_row : 0;
    M.apply_ready = false;
    M.tiles_ready = false;
}

}  // namespace axis_hip::detail
