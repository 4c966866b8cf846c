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
//   2. centroid kernels     float (lon, lat) centroids of the wet sources and
//                           the dry rows (compacted by exclusive scans); a
//                           cell whose device-libm result could round to
//                           another float than host libm's (inside an error
//                           band around a float rounding midpoint, or near
//                           the lon = 0 branch cut) is listed and recomputed
//                           on the host, so the floats are host libm's
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
constexpr int kStack = 32;
// Rows longer than this (polar rows of a conservative matrix run to ~1600
// nonzeros) are summed and compacted by a whole wave; shorter rows by one
// thread, in CSR order.
constexpr int kShortRow = 32;
constexpr int kLongWaves = 1024;

// ints[0]: largest wet count; ints[3]: number of long rows (listed in long_rows).
__global__ void wet_rows_kernel(const index_t *__restrict__ row_ptr, const index_t *__restrict__ col, const double *__restrict__ val,
                                const std::int32_t *__restrict__ mask, index_t n_rows, std::int32_t *__restrict__ wet_cnt,
                                double *__restrict__ wet_sum, int *__restrict__ ints, std::int32_t *__restrict__ long_rows) {
    const index_t r = static_cast<index_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (r >= n_rows) return;
    const index_t b = row_ptr[r], e = row_ptr[r + 1];
    if (e - b > kShortRow) {
        long_rows[atomicAdd(ints + 3, 1)] = static_cast<std::int32_t>(r);
        return;
    }
    double s = 0.0;
    std::int32_t c = 0;
    for (index_t k = b; k < e; ++k) {
        if (mask[col[k]] > 0) {
            s += val[k];
            ++c;
        }
    }
    wet_cnt[r] = c;
    wet_sum[r] = s;
    if (c > 0) atomicMax(ints, c);
}

// One wave per long row: lane-strided partial sums, then a fixed butterfly
// (deterministic; the order differs from CSR order only in rounding).
__global__ void __launch_bounds__(256) wet_long_rows_kernel(const index_t *__restrict__ row_ptr, const index_t *__restrict__ col,
                                                            const double *__restrict__ val, const std::int32_t *__restrict__ mask,
                                                            const std::int32_t *__restrict__ long_rows, int *ints,
                                                            std::int32_t *__restrict__ wet_cnt, double *__restrict__ wet_sum) {
    const int lane = threadIdx.x & 63;
    const int n_long = ints[3];
    for (int i = (blockIdx.x * blockDim.x + threadIdx.x) >> 6; i < n_long; i += (gridDim.x * blockDim.x) >> 6) {
        const index_t r = long_rows[i];
        double s = 0.0;
        int c = 0;
        for (index_t k = row_ptr[r] + lane; k < row_ptr[r + 1]; k += 64) {
            if (mask[col[k]] > 0) {
                s += val[k];
                ++c;
            }
        }
        for (int d = 32; d >= 1; d >>= 1) {
            s += __shfl_xor(s, d, 64);
            c += __shfl_xor(c, d, 64);
        }
        if (lane == 0) {
            wet_cnt[r] = c;
            wet_sum[r] = s;
            if (c > 0) atomicMax(ints, c);
        }
    }
}

// Selection flags for the compaction scans: wet sources (mask != 0) and dry
// rows (no wet weight).
__global__ void select_flags_kernel(const std::int32_t *__restrict__ v, index_t n, int dry, index_t *__restrict__ flag) {
    const index_t i = static_cast<index_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) flag[i] = dry ? (v[i] == 0) : (v[i] != 0);
}

struct CellGeom {
    const double *coords;  // [2 * n_nodes] lon block then lat block
    index_t n_nodes;
    const index_t *off, *idx;
    int cs;  // 0 SphericalDeg, 1 SphericalRad, 2 Cartesian3D
};

constexpr int kFixCap = 8192;

// Cells whose device centroid may differ from host libm's after rounding to
// float: (output position, cell) pairs, appended atomically.
struct FixList {
    std::int32_t *entry;  // [kFixCap][2]
    int *count;
};

// True when d lies within e of a midpoint between (float)d and a float
// neighbour: a value within e of d could then round to another float.  The
// midpoints of two adjacent floats are exact in double.
__device__ inline bool near_float_midpoint(double d, double e) {
    const float f = static_cast<float>(d);
    if (!isfinite(f)) return true;
    const double fd = f, lo = nextafterf(f, -INFINITY), hi = nextafterf(f, INFINITY);
    return fabs(d - 0.5 * (fd + lo)) <= e || fabs(d - 0.5 * (fd + hi)) <= e;
}

// compute_cell_centroids_xy (weight_generator.cpp:74-138) with device libm,
// rounded to float.  amb is set when host libm's result might round to
// another float.  Error budget (radians, generous): each vertex term differs
// by a few units in the last place between the two libms, so the averaged
// vector by well under (8 nv + 32) * 2^-52 per component; asin(z) and
// atan2(y, x) amplify that by at most 1 / r, r = |(x, y)| (1 - z^2 >= r^2 for
// a mean of unit vectors); degree conversion scales it by 180 / pi < 64; the
// final roundings add a few ulps of the result.  A longitude whose angle is
// within the budget of 0 may take the other side of the 2 pi shift.
__device__ inline float2 centroid_dev(const CellGeom &g, index_t c, bool &amb) {
    const double pi = 3.14159265358979323846;
    const double *lon = g.coords, *lat = g.coords + g.n_nodes;
    const index_t b = g.off[c], e = g.off[c + 1];
    const double inv = (e > b) ? 1.0 / static_cast<double>(e - b) : 0.0;
    if (g.cs == 2) {  // + and * only: the host's bits
        double sx = 0.0, sy = 0.0;
        for (index_t k = b; k < e; ++k) {
            sx += lon[g.idx[k]];
            sy += lat[g.idx[k]];
        }
        amb = false;
        return make_float2(static_cast<float>(sx * inv), static_cast<float>(sy * inv));
    }
    const bool deg = g.cs == 0;
    double sx = 0.0, sy = 0.0, sz = 0.0;
    for (index_t k = b; k < e; ++k) {
        double lo = lon[g.idx[k]], la = lat[g.idx[k]];
        if (deg) {
            lo = lo * pi / 180.0;
            la = la * pi / 180.0;
        }
        double so, co, sl, cl;
        sincos(lo, &so, &co);
        sincos(la, &sl, &cl);
        sx += cl * co;
        sy += cl * so;
        sz += sl;
    }
    const double ax = sx * inv, ay = sy * inv, az = sz * inv;
    double lat_v = asin(az);
    double lon_v = atan2(ay, ax);
    const double r = sqrt(ax * ax + ay * ay);
    const double e_rad = (8.0 * static_cast<double>(e - b) + 32.0) * 0x1p-52 * (1.0 + 1.0 / r);
    bool a = !(fabs(lon_v) > e_rad);
    if (lon_v < 0.0) lon_v += 2.0 * pi;
    double e_lon = e_rad, e_lat = e_rad;
    if (deg) {
        lon_v = lon_v * 180.0 / pi;
        lat_v = lat_v * 180.0 / pi;
        e_lon *= 64.0;
        e_lat *= 64.0;
    }
    e_lon += 16.0 * 0x1p-52 * fmax(fabs(lon_v), 1.0);
    e_lat += 16.0 * 0x1p-52 * fmax(fabs(lat_v), 1.0);
    amb = a || near_float_midpoint(lon_v, e_lon) || near_float_midpoint(lat_v, e_lat);
    return make_float2(static_cast<float>(lon_v), static_cast<float>(lat_v));
}

// KIND 0: cells with sel != 0 (wet sources); 1: cells with sel == 0 (dry
// rows); 2: every cell (pos null: position = cell).
template <int KIND>
__global__ void centroid_select_kernel(CellGeom g, index_t n, const std::int32_t *__restrict__ sel, const index_t *__restrict__ pos,
                                       float2 *__restrict__ out, std::int32_t *__restrict__ out_id, FixList fl) {
    const index_t i = static_cast<index_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (KIND == 0 && sel[i] == 0) return;
    if (KIND == 1 && sel[i] != 0) return;
    const index_t q = KIND == 2 ? i : pos[i];
    bool amb = false;
    out[q] = centroid_dev(g, i, amb);
    if (out_id) out_id[q] = static_cast<std::int32_t>(i);
    if (amb) {
        const int k = atomicAdd(fl.count, 1);
        if (k < kFixCap) {
            fl.entry[2 * k] = static_cast<std::int32_t>(q);
            fl.entry[2 * k + 1] = static_cast<std::int32_t>(i);
        }
    }
}

__global__ void scatter_fix_kernel(const std::int32_t *__restrict__ pos, const float2 *__restrict__ v, int n, float2 *__restrict__ out) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) out[pos[t]] = v[t];
}

__device__ inline float4 box_union(float4 a, float4 b) {
    return make_float4(fminf(a.x, b.x), fminf(a.y, b.y), fmaxf(a.z, b.z), fmaxf(a.w, b.w));
}

constexpr int kBoundBlocks = 256;

// (xmin, ymin, xmax, ymax) of the points: per-block partials, then one
// block folds them (launch with 256 threads).
__global__ void __launch_bounds__(256) bounds_kernel(const float2 *__restrict__ p, index_t n, float4 *__restrict__ part) {
    __shared__ float4 red[256];
    float4 b = make_float4(INFINITY, INFINITY, -INFINITY, -INFINITY);
    for (index_t q = static_cast<index_t>(blockIdx.x) * 256 + threadIdx.x; q < n; q += static_cast<index_t>(gridDim.x) * 256) {
        const float2 v = p[q];
        b = box_union(b, make_float4(v.x, v.y, v.x, v.y));
    }
    red[threadIdx.x] = b;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (static_cast<int>(threadIdx.x) < w) red[threadIdx.x] = box_union(red[threadIdx.x], red[threadIdx.x + w]);
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

__global__ void __launch_bounds__(256) bounds_fold_kernel(const float4 *__restrict__ part, float4 *__restrict__ out) {
    __shared__ float4 red[kBoundBlocks];
    red[threadIdx.x] = part[threadIdx.x];
    __syncthreads();
    for (int w = kBoundBlocks / 2; w > 0; w >>= 1) {
        if (static_cast<int>(threadIdx.x) < w) red[threadIdx.x] = box_union(red[threadIdx.x], red[threadIdx.x + w]);
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = red[0];
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

// Morton order only groups the points; the boxes come from the points
// themselves, so any quantisation is correct.
__global__ void point_keys_kernel(const float2 *__restrict__ p, index_t n, const float4 *__restrict__ bounds, std::uint32_t *__restrict__ key,
                                  std::uint32_t *__restrict__ pos) {
    const index_t q = static_cast<index_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (q >= n) return;
    const float4 bb = bounds[0];
    float sx = 65535.0f / (bb.z - bb.x), sy = 65535.0f / (bb.w - bb.y);
    if (!isfinite(sx)) sx = 0.0f;
    if (!isfinite(sy)) sy = 0.0f;
    const float x0 = isfinite(bb.x) ? bb.x : 0.0f, y0 = isfinite(bb.y) ? bb.y : 0.0f;
    const float fx = fminf(fmaxf((p[q].x - x0) * sx, 0.0f), 65535.0f);
    const float fy = fminf(fmaxf((p[q].y - y0) * sy, 0.0f), 65535.0f);
    key[q] = spread16(static_cast<std::uint32_t>(fx)) | (spread16(static_cast<std::uint32_t>(fy)) << 1);
    pos[q] = static_cast<std::uint32_t>(q);
}

// Sorted points, padded to n_pad (a whole number of leaves) with NaN
// points that never win.
__global__ void gather_points_kernel(const std::uint32_t *__restrict__ pos, const float2 *__restrict__ p, const std::int32_t *__restrict__ pidx,
                                     index_t n, index_t n_pad, float2 *__restrict__ pts, std::int32_t *__restrict__ idx) {
    const index_t q = static_cast<index_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (q >= n_pad) return;
    if (q >= n) {
        pts[q] = make_float2(NAN, NAN);
        idx[q] = std::numeric_limits<std::int32_t>::max();
        return;
    }
    const std::uint32_t k = pos[q];
    pts[q] = p[k];
    idx[q] = pidx[k];
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

// One thread per query; the traversal stack sits in LDS ([depth][thread],
// conflict-free): a stack in scratch memory costs a memory round trip per
// push and pop.  The tree height is at most 26 (2^26 leaves of 32 points), so
// kStack entries always suffice.  Leaves are padded to 32 points with NaN
// (never nearer than anything) and read as float4 pairs.
constexpr int kNearThreads = 128;

__global__ void __launch_bounds__(kNearThreads) nearest_kernel(const float4 *__restrict__ box, const float4 *__restrict__ pts2,
                                                               const std::int32_t *__restrict__ idx, index_t n_leaves, index_t P,
                                                               const float2 *__restrict__ query, const std::int32_t *__restrict__ qrow,
                                                               const std::uint32_t *__restrict__ qorder, index_t n_q,
                                                               std::int32_t *__restrict__ nearest) {
    __shared__ std::int32_t stack[kStack * kNearThreads];
    const index_t tt = static_cast<index_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (tt >= n_q) return;
    const index_t t = qorder[tt];  // queries in Morton order: a wave walks one neighbourhood
    const float qx = query[t].x, qy = query[t].y;
    float best = INFINITY;
    std::int32_t bi = std::numeric_limits<std::int32_t>::max();
    std::int32_t *st = stack + threadIdx.x;
    int sp = 0;
    st[0] = 1;
    sp = 1;
    while (sp > 0) {
        const std::int32_t node = st[--sp * kNearThreads];
        if (node >= P) {
            const index_t l = node - P;
            if (l >= n_leaves) continue;
            const float4 *lp = pts2 + l * (kLeafPts / 2);
            const std::int32_t *li = idx + l * kLeafPts;
#pragma unroll 4
            for (int j = 0; j < kLeafPts / 2; ++j) {
                const float4 v = lp[j];
                float dx = v.x - qx, dy = v.y - qy;
                float d2 = dx * dx + dy * dy;
                if (d2 < best || (d2 == best && li[2 * j] < bi)) {
                    best = d2;
                    bi = li[2 * j];
                }
                dx = v.z - qx;
                dy = v.w - qy;
                d2 = dx * dx + dy * dy;
                if (d2 < best || (d2 == best && li[2 * j + 1] < bi)) {
                    best = d2;
                    bi = li[2 * j + 1];
                }
            }
            continue;
        }
        const std::int32_t c0 = 2 * node, c1 = 2 * node + 1;
        const float d0 = box_d2(box[c0], qx, qy), d1 = box_d2(box[c1], qx, qy);
        // nearer child on top of the stack
        const std::int32_t near_c = d0 <= d1 ? c0 : c1, far_c = d0 <= d1 ? c1 : c0;
        const float near_d = fminf(d0, d1), far_d = fmaxf(d0, d1);
        if (!(far_d > best) && sp < kStack) st[sp++ * kNearThreads] = far_c;
        if (!(near_d > best) && sp < kStack) st[sp++ * kNearThreads] = near_c;
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

// Short rows (and every dry row) by one thread each; long rows with wet
// weights are left to coastal_fill_long_kernel.
__global__ void coastal_fill_kernel(const index_t *__restrict__ row_ptr, const index_t *__restrict__ col, const double *__restrict__ val,
                                    const std::int32_t *__restrict__ mask, const std::int32_t *__restrict__ wet_cnt,
                                    const double *__restrict__ wet_sum, const std::int32_t *__restrict__ nearest, index_t n_rows,
                                    const index_t *__restrict__ new_ptr, index_t *__restrict__ new_col, double *__restrict__ new_val) {
    const index_t r = static_cast<index_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (r >= n_rows) return;
    index_t o = new_ptr[r];
    const index_t b = row_ptr[r], e = row_ptr[r + 1];
    if (wet_cnt[r] > 0) {
        if (e - b > kShortRow) return;
        const double s = wet_sum[r];
        const bool renorm = s > 0.0 && s < 1.0;
        for (index_t k = b; k < e; ++k) {
            const index_t c = col[k];
            if (mask[c] > 0) {
                double w = val[k];
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

// One wave per long row: 64 entries at a time, wet ones placed in CSR order
// by ballot prefix counts.
__global__ void __launch_bounds__(256) coastal_fill_long_kernel(const index_t *__restrict__ row_ptr, const index_t *__restrict__ col,
                                                                const double *__restrict__ val, const std::int32_t *__restrict__ mask,
                                                                const std::int32_t *__restrict__ long_rows, const int *__restrict__ ints,
                                                                const std::int32_t *__restrict__ wet_cnt, const double *__restrict__ wet_sum,
                                                                const index_t *__restrict__ new_ptr, index_t *__restrict__ new_col,
                                                                double *__restrict__ new_val) {
    const int lane = threadIdx.x & 63;
    const unsigned long long below = lane == 0 ? 0ull : (~0ull >> (64 - lane));
    const int n_long = ints[3];
    for (int i = (blockIdx.x * blockDim.x + threadIdx.x) >> 6; i < n_long; i += (gridDim.x * blockDim.x) >> 6) {
        const index_t r = long_rows[i];
        if (wet_cnt[r] == 0) continue;
        const double s = wet_sum[r];
        const bool renorm = s > 0.0 && s < 1.0;
        index_t o = new_ptr[r];
        const index_t e = row_ptr[r + 1];
        for (index_t base = row_ptr[r]; base < e; base += 64) {
            const index_t k = base + lane;
            index_t c = 0;
            bool wet = false;
            if (k < e) {
                c = col[k];
                wet = mask[c] > 0;
            }
            const unsigned long long m = __ballot(wet);
            if (wet) {
                double w = val[k];
                if (renorm) w /= s;
                const index_t at = o + __popcll(m & below);
                new_col[at] = c;
                new_val[at] = w;
            }
            o += __popcll(m);
        }
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

CellGeom geom_of(const Mesh &m) {
    const int cs = m.coord_system() == CoordinateSystem::SphericalDeg ? 0 : (m.coord_system() == CoordinateSystem::SphericalRad ? 1 : 2);
    return CellGeom{m.d_coords.get(), m.n_nodes(), m.d_offsets.get(), m.d_indices.get(), cs};
}

// Recomputes the listed cells' centroids with host libm and writes them
// over the device results.  count > kFixCap (never expected): every entry of
// out is recomputed from ids (the cell of each position; null: position =
// cell).
void fix_centroids(const Mesh &m, const std::int32_t *d_entry, int count, float2 *out, const std::int32_t *ids, index_t n_out, hipStream_t s) {
    if (count <= 0) return;
    if (count <= kFixCap) {
        std::vector<std::int32_t> e(2 * static_cast<std::size_t>(count));
        AXIS_HIP_CHECK(hipMemcpyAsync(e.data(), d_entry, e.size() * sizeof(std::int32_t), hipMemcpyDeviceToHost, s));
        AXIS_HIP_CHECK(hipStreamSynchronize(s));
        std::vector<std::int32_t> pos(count);
        std::vector<float2> v(count);
        SinCosMemo memo;
        for (int k = 0; k < count; ++k) {
            double cx, cy;
            centroid_lonlat(m, e[2 * k + 1], memo, cx, cy);
            pos[k] = e[2 * k];
            v[k] = make_float2(static_cast<float>(cx), static_cast<float>(cy));
        }
        DeviceBuffer<std::int32_t> d_pos(count);
        DeviceBuffer<float2> d_v(count);
        d_pos.upload(pos.data(), count, s);
        d_v.upload(v.data(), count, s);
        scatter_fix_kernel<<<grid_for(count, 256), 256, 0, s>>>(d_pos.get(), d_v.get(), count, out);
        AXIS_LAUNCH_CHECK();
        AXIS_HIP_CHECK(hipStreamSynchronize(s));
        return;
    }
    std::vector<std::int32_t> id(n_out);
    if (ids) {
        AXIS_HIP_CHECK(hipMemcpyAsync(id.data(), ids, n_out * sizeof(std::int32_t), hipMemcpyDeviceToHost, s));
        AXIS_HIP_CHECK(hipStreamSynchronize(s));
    } else {
        for (index_t q = 0; q < n_out; ++q) id[q] = static_cast<std::int32_t>(q);
    }
    std::vector<float2> v(n_out);
    parallel_range(n_out, [&](index_t b, index_t e) {
        SinCosMemo memo;
        for (index_t q = b; q < e; ++q) {
            double cx, cy;
            centroid_lonlat(m, id[q], memo, cx, cy);
            v[q] = make_float2(static_cast<float>(cx), static_cast<float>(cy));
        }
    });
    AXIS_HIP_CHECK(hipMemcpyAsync(out, v.data(), n_out * sizeof(float2), hipMemcpyHostToDevice, s));
    AXIS_HIP_CHECK(hipStreamSynchronize(s));
}

}  // namespace

index_t cell_centroids_f32(const Mesh &m, float2 *out, hipStream_t s) {
    const index_t n = m.n_cells();
    if (n == 0) return 0;
    DeviceBuffer<std::int32_t> fix_entry(2 * kFixCap), fix_count(1);
    fill_int_kernel<<<1, 64, 0, s>>>(fix_count.get(), 1, 0);
    AXIS_LAUNCH_CHECK();
    centroid_select_kernel<2><<<grid_for(n, 256), 256, 0, s>>>(geom_of(m), n, nullptr, nullptr, out, nullptr,
                                                               FixList{fix_entry.get(), fix_count.get()});
    AXIS_LAUNCH_CHECK();
    int count = 0;
    AXIS_HIP_CHECK(hipMemcpyAsync(&count, fix_count.get(), sizeof(int), hipMemcpyDeviceToHost, s));
    AXIS_HIP_CHECK(hipStreamSynchronize(s));
    fix_centroids(m, fix_entry.get(), count, out, nullptr, n, s);
    return count;
}

void coastal_pass(const Mesh &src, const Mesh &dst, ExtrapolationAction extrap, InterpolationMatrix &M, hipStream_t s) {
    if (!src.has_mask()) return;
    const index_t n_dst = M.n_dst, n_src = M.n_src;
    if (n_dst == 0) return;
    PhaseTrace tr("coastal", s);
    const index_t nr = std::max<index_t>(n_dst, 1), ns = std::max<index_t>(n_src, 1);
    // ints: largest wet count, fix counts (sources, queries), long-row count
    DeviceBuffer<std::int32_t> wet_cnt(nr), nearest(nr), ints(4), long_rows(nr);
    DeviceBuffer<double> wet_sum(nr);
    fill_int_kernel<<<grid_for(n_dst, 256), 256, 0, s>>>(nearest.get(), n_dst, -1);
    AXIS_LAUNCH_CHECK();
    fill_int_kernel<<<1, 64, 0, s>>>(ints.get(), 4, 0);
    AXIS_LAUNCH_CHECK();
    wet_rows_kernel<<<grid_for(n_dst, 256), 256, 0, s>>>(M.row_ptr.get(), M.col.get(), M.val.get(), src.d_mask.get(), n_dst, wet_cnt.get(),
                                                         wet_sum.get(), ints.get(), long_rows.get());
    AXIS_LAUNCH_CHECK();
    wet_long_rows_kernel<<<kLongWaves / 4, 256, 0, s>>>(M.row_ptr.get(), M.col.get(), M.val.get(), src.d_mask.get(), long_rows.get(), ints.get(),
                                                        wet_cnt.get(), wet_sum.get());
    AXIS_LAUNCH_CHECK();
    int h_ints[4] = {0, 0, 0, 0};
    if (extrap == ExtrapolationAction::NearestWet) {
        // compaction of the wet sources and the dry rows, with their centroids
        DeviceBuffer<index_t> flag_s(ns + 1), pos_s(ns + 1), flag_d(nr + 1), pos_d(nr + 1);
        select_flags_kernel<<<grid_for(n_src, 256), 256, 0, s>>>(src.d_mask.get(), n_src, 0, flag_s.get());
        AXIS_LAUNCH_CHECK();
        exclusive_scan_i64(flag_s.get(), pos_s.get(), n_src, s);
        select_flags_kernel<<<grid_for(n_dst, 256), 256, 0, s>>>(wet_cnt.get(), n_dst, 1, flag_d.get());
        AXIS_LAUNCH_CHECK();
        exclusive_scan_i64(flag_d.get(), pos_d.get(), n_dst, s);
        DeviceBuffer<float2> p0(ns), q(nr);
        DeviceBuffer<std::int32_t> pid0(ns), qrow(nr), fix_entry(4 * kFixCap);
        centroid_select_kernel<0><<<grid_for(n_src, 256), 256, 0, s>>>(geom_of(src), n_src, src.d_mask.get(), pos_s.get(), p0.get(), pid0.get(),
                                                                       FixList{fix_entry.get(), ints.get() + 1});
        AXIS_LAUNCH_CHECK();
        centroid_select_kernel<1><<<grid_for(n_dst, 256), 256, 0, s>>>(geom_of(dst), n_dst, wet_cnt.get(), pos_d.get(), q.get(), qrow.get(),
                                                                       FixList{fix_entry.get() + 2 * kFixCap, ints.get() + 2});
        AXIS_LAUNCH_CHECK();
        index_t n_wet = 0, n_q = 0;
        AXIS_HIP_CHECK(hipMemcpyAsync(&n_wet, pos_s.get() + n_src, sizeof(index_t), hipMemcpyDeviceToHost, s));
        AXIS_HIP_CHECK(hipMemcpyAsync(&n_q, pos_d.get() + n_dst, sizeof(index_t), hipMemcpyDeviceToHost, s));
        AXIS_HIP_CHECK(hipMemcpyAsync(h_ints, ints.get(), sizeof h_ints, hipMemcpyDeviceToHost, s));
        AXIS_HIP_CHECK(hipStreamSynchronize(s));
        tr.mark("wet_rows+centroids");
        if (n_wet > 0 && n_q > 0) {
            fix_centroids(src, fix_entry.get(), h_ints[1], p0.get(), pid0.get(), n_wet, s);
            fix_centroids(dst, fix_entry.get() + 2 * kFixCap, h_ints[2], q.get(), qrow.get(), n_q, s);
            tr.mark("host_fixups");
            DeviceBuffer<float4> part(kBoundBlocks), bounds(1);
            bounds_kernel<<<kBoundBlocks, 256, 0, s>>>(p0.get(), n_wet, part.get());
            AXIS_LAUNCH_CHECK();
            bounds_fold_kernel<<<1, kBoundBlocks, 0, s>>>(part.get(), bounds.get());
            AXIS_LAUNCH_CHECK();
            DeviceBuffer<std::uint32_t> key(n_wet), kpos(n_wet);
            point_keys_kernel<<<grid_for(n_wet, 256), 256, 0, s>>>(p0.get(), n_wet, bounds.get(), key.get(), kpos.get());
            AXIS_LAUNCH_CHECK();
            radix_sort_pairs_u32(key.get(), kpos.get(), static_cast<std::size_t>(n_wet), 32, s);
            const index_t n_leaves = (n_wet + kLeafPts - 1) / kLeafPts, n_pad = n_leaves * kLeafPts;
            index_t P = 1;
            while (P < n_leaves) P *= 2;
            if (P > (index_t(1) << 26)) throw InvalidArgument("coastal pass: more than 2^31 wet source cells");
            DeviceBuffer<float2> pts(n_pad);
            DeviceBuffer<std::int32_t> pidx(n_pad);
            gather_points_kernel<<<grid_for(n_pad, 256), 256, 0, s>>>(kpos.get(), p0.get(), pid0.get(), n_wet, n_pad, pts.get(), pidx.get());
            AXIS_LAUNCH_CHECK();
            DeviceBuffer<float4> box(2 * P);
            leaf_boxes_kernel<<<grid_for(P, 256), 256, 0, s>>>(pts.get(), n_wet, n_leaves, P, box.get());
            AXIS_LAUNCH_CHECK();
            for (index_t lo = P / 2; lo >= 1; lo /= 2) {
                inner_boxes_kernel<<<grid_for(lo, 256), 256, 0, s>>>(box.get(), lo);
                AXIS_LAUNCH_CHECK();
            }
            DeviceBuffer<std::uint32_t> qkey(n_q), qorder(n_q);
            point_keys_kernel<<<grid_for(n_q, 256), 256, 0, s>>>(q.get(), n_q, bounds.get(), qkey.get(), qorder.get());
            AXIS_LAUNCH_CHECK();
            radix_sort_pairs_u32(qkey.get(), qorder.get(), static_cast<std::size_t>(n_q), 32, s);
            nearest_kernel<<<grid_for(n_q, kNearThreads), kNearThreads, 0, s>>>(box.get(), reinterpret_cast<const float4 *>(pts.get()), pidx.get(),
                                                                                n_leaves, P, q.get(), qrow.get(), qorder.get(), n_q, nearest.get());
            AXIS_LAUNCH_CHECK();
            AXIS_HIP_CHECK(hipStreamSynchronize(s));  // the buffers above are released on scope exit
            tr.mark("nearest");
        }
    }
    DeviceBuffer<index_t> cnt(nr + 1), new_ptr(nr + 1);
    new_counts_kernel<<<grid_for(n_dst, 256), 256, 0, s>>>(wet_cnt.get(), nearest.get(), n_dst, cnt.get());
    AXIS_LAUNCH_CHECK();
    exclusive_scan_i64(cnt.get(), new_ptr.get(), n_dst, s);
    index_t new_nnz = 0;
    AXIS_HIP_CHECK(hipMemcpyAsync(&new_nnz, new_ptr.get() + n_dst, sizeof(index_t), hipMemcpyDeviceToHost, s));
    AXIS_HIP_CHECK(hipMemcpyAsync(h_ints, ints.get(), sizeof(int), hipMemcpyDeviceToHost, s));
    AXIS_HIP_CHECK(hipStreamSynchronize(s));
    DeviceBuffer<index_t> new_col(std::max<index_t>(new_nnz, 1));
    DeviceBuffer<double> new_val(std::max<index_t>(new_nnz, 1));
    coastal_fill_kernel<<<grid_for(n_dst, 256), 256, 0, s>>>(M.row_ptr.get(), M.col.get(), M.val.get(), src.d_mask.get(), wet_cnt.get(),
                                                             wet_sum.get(), nearest.get(), n_dst, new_ptr.get(), new_col.get(), new_val.get());
    AXIS_LAUNCH_CHECK();
    coastal_fill_long_kernel<<<kLongWaves / 4, 256, 0, s>>>(M.row_ptr.get(), M.col.get(), M.val.get(), src.d_mask.get(), long_rows.get(), ints.get(),
                                                            wet_cnt.get(), wet_sum.get(), new_ptr.get(), new_col.get(), new_val.get());
    AXIS_LAUNCH_CHECK();
    AXIS_HIP_CHECK(hipStreamSynchronize(s));  // the old CSR and the work buffers are released below
    tr.mark("assemble");
    M.row_ptr = std::move(new_ptr);
    M.col = std::move(new_col);
    M.val = std::move(new_val);
    M.nnz = new_nnz;
    M.max_row_nnz = new_nnz > 0 ? std::max(h_ints[0], 1) : 0;
    M.apply_ready = false;
    M.tiles_ready = false;
}

}  // namespace axis_hip::detail
