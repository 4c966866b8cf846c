// Nearest-point searches of weight generation, on the GPU:
//
//   * NearestNeighbor weights (WeightGenerator::generate_nearest,
//     weight_generator.cpp:2025-2271): for regular source and destination
//     lat-lon grids the analytic index rounding of generate_nearest_rect;
//     otherwise each destination cell maps with weight 1 to the source cell
//     whose centroid point is nearest -- the float unit vector of the
//     (lon, lat) centroid on a sphere, the float centroid in the plane.
//   * The masked-source post-pass (coastal_renormalize_and_extrapolate,
//     weight_generator.cpp:1736-1881): with a source cell mask, only weights
//     on wet (mask > 0) sources are kept, a row's kept weights divided by
//     their sum when it lies in (0, 1); with ExtrapolationAction::NearestWet
//     a row left without weights gets weight 1 on the wet (mask != 0) source
//     whose float (lon, lat) centroid is nearest in the plane.  Fractions,
//     areas and the rows' spatial keys are unchanged (:1879).
//
// Shared machinery:
//   * centroid kernels: compute_cell_centroids_xy (:74-138) with device libm;
//     a cell whose result could round to another float than host libm's
//     (inside an error band around a float rounding midpoint, or at the
//     lon = 0 branch cut) is listed and recomputed on the host, so the float
//     points equal the reference's host-computed ones bit for bit;
//   * exact nearest search over float points (2D or 3D): Morton sort, leaves
//     of 4 points, an implicit complete binary tree of float boxes (heap
//     order, root 1), one thread per query with its traversal stack in LDS;
//     smallest float squared distance wins, ties go to the smallest source
//     index (ArborX leaves exact ties open).
#include <algorithm>
#include <chrono>
#include <cmath>
#include <limits>

#include "axis_hip.hpp"
#include "host_util.hpp"

namespace axis_hip::detail {
namespace {

constexpr int kLeafPts = 4;  // 32: 7.6 ms, 16: 5.5, 8: 4.2, 4: 3.5, 2: 3.3 for 1.04 M queries against 885 K unit vectors
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

// ------------------------------------------------------------ point types
// D = 2: (lon, lat) points as float2, node boxes as one float4
// (xmin, ymin, xmax, ymax).  D = 3: unit vectors as float4 (x, y, z, 0),
// node boxes as two float4 (lo, hi).  Squared distances are summed x, y, z in
// that order with separate multiplies and adds, so the 2D sums equal the
// reference's Point<2> ones and the 3D sums its Point<3> ones.
template <int D>
struct PT;

template <>
struct PT<2> {
    using P = float2;
    struct B {
        float4 v;
    };
    static constexpr int kKeyBits = 12;  // 24-bit keys: three radix passes
    __device__ static B empty() { return B{make_float4(INFINITY, INFINITY, -INFINITY, -INFINITY)}; }
    __device__ static B of(P p) { return B{make_float4(p.x, p.y, p.x, p.y)}; }
    __device__ static B join(B a, B b) { return B{make_float4(fminf(a.v.x, b.v.x), fminf(a.v.y, b.v.y), fmaxf(a.v.z, b.v.z), fmaxf(a.v.w, b.v.w))}; }
    __device__ static float lo(const B &b, int a) { return a == 0 ? b.v.x : b.v.y; }
    __device__ static float hi(const B &b, int a) { return a == 0 ? b.v.z : b.v.w; }
    __device__ static float c(const P &p, int a) { return a == 0 ? p.x : p.y; }
    __device__ static P nan_point() { return make_float2(NAN, NAN); }
    __device__ static float d2(const P &p, const P &q) {
        const float dx = p.x - q.x, dy = p.y - q.y;
        return dx * dx + dy * dy;
    }
    // For any point inside the box the rounded differences and squares are
    // monotone in the exact ones, so this never exceeds the float squared
    // distance to a point inside: pruning on "box > best" keeps every point
    // that could win or tie.
    __device__ static float bd2(const B &b, const P &q) {
        const float dx = fmaxf(fmaxf(b.v.x - q.x, q.x - b.v.z), 0.0f);
        const float dy = fmaxf(fmaxf(b.v.y - q.y, q.y - b.v.w), 0.0f);
        return dx * dx + dy * dy;
    }
};

template <>
struct PT<3> {
    using P = float4;
    struct B {
        float4 l, h;
    };
    static constexpr int kKeyBits = 8;  // 24-bit keys: three radix passes
    __device__ static B empty() { return B{make_float4(INFINITY, INFINITY, INFINITY, 0.0f), make_float4(-INFINITY, -INFINITY, -INFINITY, 0.0f)}; }
    __device__ static B of(P p) { return B{p, p}; }
    __device__ static B join(B a, B b) {
        return B{make_float4(fminf(a.l.x, b.l.x), fminf(a.l.y, b.l.y), fminf(a.l.z, b.l.z), 0.0f),
                 make_float4(fmaxf(a.h.x, b.h.x), fmaxf(a.h.y, b.h.y), fmaxf(a.h.z, b.h.z), 0.0f)};
    }
    __device__ static float lo(const B &b, int a) { return a == 0 ? b.l.x : (a == 1 ? b.l.y : b.l.z); }
    __device__ static float hi(const B &b, int a) { return a == 0 ? b.h.x : (a == 1 ? b.h.y : b.h.z); }
    __device__ static float c(const P &p, int a) { return a == 0 ? p.x : (a == 1 ? p.y : p.z); }
    __device__ static P nan_point() { return make_float4(NAN, NAN, NAN, 0.0f); }
    __device__ static float d2(const P &p, const P &q) {
        const float dx = p.x - q.x, dy = p.y - q.y, dz = p.z - q.z;
        return dx * dx + dy * dy + dz * dz;
    }
    __device__ static float bd2(const B &b, const P &q) {
        const float dx = fmaxf(fmaxf(b.l.x - q.x, q.x - b.h.x), 0.0f);
        const float dy = fmaxf(fmaxf(b.l.y - q.y, q.y - b.h.y), 0.0f);
        const float dz = fmaxf(fmaxf(b.l.z - q.z, q.z - b.h.z), 0.0f);
        return dx * dx + dy * dy + dz * dz;
    }
};

// ------------------------------------------------------------- centroids
constexpr int kFixCap = 8192;

// Cells whose device point may differ from host libm's after rounding to
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

// compute_cell_centroids_xy (weight_generator.cpp:74-138) with device libm:
// (lon, lat) in the mesh's unit, with bounds e_rad (radians) on its distance
// from host libm's result and cut = the longitude's side of the 2 pi shift
// is uncertain.  Error budget (generous): each vertex term differs by a few
// units in the last place between the two libms, so the averaged vector by
// well under (8 nv + 32) 2^-52 per component; asin(z) and atan2(y, x) amplify
// that by at most 1 / r, r = |(x, y)| (1 - z^2 >= r^2 for a mean of unit
// vectors).  Planar meshes use + and * only: no error.
__device__ inline void centroid_dev(const CellGeom &g, index_t c, double &lon_v, double &lat_v, double &e_rad, bool &cut) {
    const double pi = 3.14159265358979323846;
    const double *lon = g.coords, *lat = g.coords + g.n_nodes;
    const index_t b = g.off[c], e = g.off[c + 1];
    const double inv = (e > b) ? 1.0 / static_cast<double>(e - b) : 0.0;
    if (g.cs == 2) {
        double sx = 0.0, sy = 0.0;
        for (index_t k = b; k < e; ++k) {
            sx += lon[g.idx[k]];
            sy += lat[g.idx[k]];
        }
        lon_v = sx * inv;
        lat_v = sy * inv;
        e_rad = 0.0;
        cut = false;
        return;
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
    lat_v = asin(az);
    lon_v = atan2(ay, ax);
    const double r = sqrt(ax * ax + ay * ay);
    e_rad = (8.0 * static_cast<double>(e - b) + 32.0) * 0x1p-52 * (1.0 + 1.0 / r);
    cut = !(fabs(lon_v) > e_rad);  // NaN lands here too
    if (lon_v < 0.0) lon_v += 2.0 * pi;
    e_rad += 16.0 * 0x1p-52 * 2.0 * pi;  // the shift and the unit round trips below
    if (deg) {
        lon_v = lon_v * 180.0 / pi;
        lat_v = lat_v * 180.0 / pi;
    }
}

// The float point of cell c as the reference builds it, and whether host
// libm might give another float.  D = 2: the (lon, lat) centroid.  D = 3:
// the unit vector of the centroid (degrees converted as lon * pi / 180),
// whose components move by at most twice the angle error plus the libm
// differences.
template <int D>
__device__ inline typename PT<D>::P cell_point(const CellGeom &g, index_t c, bool &amb) {
    double lon_v, lat_v, e_rad;
    bool cut;
    centroid_dev(g, c, lon_v, lat_v, e_rad, cut);
    if constexpr (D == 2) {
        const double scale = g.cs == 0 ? 64.0 : 1.0;  // 180 / pi < 64
        const double e_lon = e_rad * scale + 16.0 * 0x1p-52 * fmax(fabs(lon_v), 1.0);
        const double e_lat = e_rad * scale + 16.0 * 0x1p-52 * fmax(fabs(lat_v), 1.0);
        amb = g.cs != 2 && (cut || near_float_midpoint(lon_v, e_lon) || near_float_midpoint(lat_v, e_lat));
        return make_float2(static_cast<float>(lon_v), static_cast<float>(lat_v));
    } else {
        const double pi = 3.14159265358979323846;
        double lo = lon_v, la = lat_v;
        if (g.cs == 0) {
            lo = lo * pi / 180.0;
            la = la * pi / 180.0;
        }
        double so, co, sl, cl;
        sincos(lo, &so, &co);
        sincos(la, &sl, &cl);
        const double x = cl * co, y = cl * so, z = sl;
        const double e = 2.0 * e_rad + 32.0 * 0x1p-52;
        amb = near_float_midpoint(x, e) || near_float_midpoint(y, e) || near_float_midpoint(z, e);
        return make_float4(static_cast<float>(x), static_cast<float>(y), static_cast<float>(z), 0.0f);
    }
}

// KIND 0: cells with sel != 0 (wet sources); 1: cells with sel == 0 (dry
// rows); 2: every cell (pos null: position = cell).
template <int KIND, int D>
__global__ void centroid_select_kernel(CellGeom g, index_t n, const std::int32_t *__restrict__ sel, const index_t *__restrict__ pos,
                                       typename PT<D>::P *__restrict__ out, std::int32_t *__restrict__ out_id, FixList fl) {
    const index_t i = static_cast<index_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (KIND == 0 && sel[i] == 0) return;
    if (KIND == 1 && sel[i] != 0) return;
    const index_t q = KIND == 2 ? i : pos[i];
    bool amb = false;
    out[q] = cell_point<D>(g, i, amb);
    if (out_id) out_id[q] = static_cast<std::int32_t>(i);
    if (amb) {
        const int k = atomicAdd(fl.count, 1);
        if (k < kFixCap) {
            fl.entry[2 * k] = static_cast<std::int32_t>(q);
            fl.entry[2 * k + 1] = static_cast<std::int32_t>(i);
        }
    }
}

template <class P>
__global__ void scatter_fix_kernel(const std::int32_t *__restrict__ pos, const P *__restrict__ v, int n, P *__restrict__ out) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) out[pos[t]] = v[t];
}

// ------------------------------------------------------------ point tree
constexpr int kBoundBlocks = 256;

// Bounding box of the points: per-block partials, then one block folds them.
template <int D>
__global__ void __launch_bounds__(256) bounds_kernel(const typename PT<D>::P *__restrict__ p, index_t n, typename PT<D>::B *__restrict__ part) {
    using T = PT<D>;
    __shared__ typename T::B red[256];
    typename T::B b = T::empty();
    for (index_t q = static_cast<index_t>(blockIdx.x) * 256 + threadIdx.x; q < n; q += static_cast<index_t>(gridDim.x) * 256) b = T::join(b, T::of(p[q]));
    red[threadIdx.x] = b;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (static_cast<int>(threadIdx.x) < w) red[threadIdx.x] = T::join(red[threadIdx.x], red[threadIdx.x + w]);
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

template <int D>
__global__ void __launch_bounds__(256) bounds_fold_kernel(const typename PT<D>::B *__restrict__ part, typename PT<D>::B *__restrict__ out) {
    using T = PT<D>;
    __shared__ typename T::B red[kBoundBlocks];
    red[threadIdx.x] = part[threadIdx.x];
    __syncthreads();
    for (int w = kBoundBlocks / 2; w > 0; w >>= 1) {
        if (static_cast<int>(threadIdx.x) < w) red[threadIdx.x] = T::join(red[threadIdx.x], red[threadIdx.x + w]);
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = red[0];
}

// Morton code: kKeyBits bits of each quantised coordinate, interleaved.
__device__ inline std::uint32_t spread_bits(std::uint32_t v, int d) {
    std::uint32_t r = 0;
    for (int b = 0; b < 16; ++b) r |= ((v >> b) & 1u) << (b * d);
    return r;
}

// Morton order only groups the points; the boxes come from the points
// themselves, so any quantisation is correct.
template <int D>
__global__ void point_keys_kernel(const typename PT<D>::P *__restrict__ p, index_t n, const typename PT<D>::B *__restrict__ bounds,
                                  std::uint32_t *__restrict__ key, std::uint32_t *__restrict__ pos) {
    using T = PT<D>;
    const index_t q = static_cast<index_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (q >= n) return;
    const typename T::B bb = bounds[0];
    const float top = static_cast<float>((1u << T::kKeyBits) - 1u);
    std::uint32_t k = 0;
    for (int a = 0; a < D; ++a) {
        const float l = T::lo(bb, a), h = T::hi(bb, a);
        float sc = top / (h - l);
        if (!isfinite(sc)) sc = 0.0f;
        const float x0 = isfinite(l) ? l : 0.0f;
        const float f = fminf(fmaxf((T::c(p[q], a) - x0) * sc, 0.0f), top);  // NaN -> 0
        k |= spread_bits(static_cast<std::uint32_t>(f), D) << a;
    }
    key[q] = k;
    pos[q] = static_cast<std::uint32_t>(q);
}

// Sorted points, padded to n_pad (a whole number of leaves) with NaN
// points that never win.
template <int D>
__global__ void gather_points_kernel(const std::uint32_t *__restrict__ pos, const typename PT<D>::P *__restrict__ p,
                                     const std::int32_t *__restrict__ pidx, index_t n, index_t n_pad, typename PT<D>::P *__restrict__ pts,
                                     std::int32_t *__restrict__ idx) {
    const index_t q = static_cast<index_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (q >= n_pad) return;
    if (q >= n) {
        pts[q] = PT<D>::nan_point();
        idx[q] = std::numeric_limits<std::int32_t>::max();
        return;
    }
    const std::uint32_t k = pos[q];
    pts[q] = p[k];
    idx[q] = pidx ? pidx[k] : static_cast<std::int32_t>(k);
}

// Leaf boxes (a leaf past the last one gets an empty box, never entered: the
// traversal checks the leaf range), then one launch per inner level.
template <int D>
__global__ void leaf_boxes_kernel(const typename PT<D>::P *__restrict__ pts, index_t n, index_t n_leaves, index_t P,
                                  typename PT<D>::B *__restrict__ box) {
    using T = PT<D>;
    const index_t l = static_cast<index_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (l >= P) return;
    typename T::B b = T::empty();
    if (l < n_leaves) {
        const index_t e = std::min<index_t>(n, (l + 1) * kLeafPts);
        for (index_t q = l * kLeafPts; q < e; ++q) b = T::join(b, T::of(pts[q]));
    }
    box[P + l] = b;
}

template <int D>
__global__ void inner_boxes_kernel(typename PT<D>::B *__restrict__ box, index_t lo) {
    const index_t i = lo + static_cast<index_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= 2 * lo) return;
    box[i] = PT<D>::join(box[2 * i], box[2 * i + 1]);
}

// One thread per query (optionally taken in the order qorder); the traversal
// stack sits in LDS ([depth][thread], conflict-free): a stack in scratch
// memory costs a memory round trip per push and pop.  The host refuses trees
// taller than kStack - 1 levels, so the stack never overflows.
// out[qrow ? qrow[t] : t] receives the nearest point's id (left alone when
// every distance is NaN).
constexpr int kNearThreads = 128;

template <int D>
__global__ void __launch_bounds__(kNearThreads) nearest_kernel(const typename PT<D>::B *__restrict__ box, const typename PT<D>::P *__restrict__ pts,
                                                               const std::int32_t *__restrict__ idx, index_t n_leaves, index_t P,
                                                               const typename PT<D>::P *__restrict__ query, const std::int32_t *__restrict__ qrow,
                                                               const std::uint32_t *__restrict__ qorder, index_t n_q,
                                                               std::int32_t *__restrict__ out) {
    using T = PT<D>;
    __shared__ std::int32_t stack[kStack * kNearThreads];
    const index_t tt = static_cast<index_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (tt >= n_q) return;
    const index_t t = qorder ? static_cast<index_t>(qorder[tt]) : tt;
    const typename T::P qp = query[t];
    float best = INFINITY;
    std::int32_t bi = std::numeric_limits<std::int32_t>::max();
    std::int32_t *st = stack + threadIdx.x;
    int sp = 1;
    st[0] = 1;
    while (sp > 0) {
        const std::int32_t node = st[--sp * kNearThreads];
        if (node >= P) {
            const index_t l = node - P;
            if (l >= n_leaves) continue;
            const typename T::P *lp = pts + l * kLeafPts;
            const std::int32_t *li = idx + l * kLeafPts;
#pragma unroll 4
            for (int j = 0; j < kLeafPts; ++j) {
                const float d2 = T::d2(lp[j], qp);
                if (d2 < best || (d2 == best && li[j] < bi)) {
                    best = d2;
                    bi = li[j];
                }
            }
            continue;
        }
        const std::int32_t c0 = 2 * node, c1 = 2 * node + 1;
        const float d0 = T::bd2(box[c0], qp), d1 = T::bd2(box[c1], qp);
        // nearer child on top of the stack
        const std::int32_t near_c = d0 <= d1 ? c0 : c1, far_c = d0 <= d1 ? c1 : c0;
        const float near_d = fminf(d0, d1), far_d = fmaxf(d0, d1);
        if (!(far_d > best) && sp < kStack) st[sp++ * kNearThreads] = far_c;
        if (!(near_d > best) && sp < kStack) st[sp++ * kNearThreads] = near_c;
    }
    if (bi != std::numeric_limits<std::int32_t>::max()) out[qrow ? static_cast<index_t>(qrow[t]) : t] = bi;
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

// ---------------------------------------------------- NearestNeighbor kernels
// generate_nearest_rect (weight_generator.cpp:2025-2083): dst cell j = (i, jj)
// with i = j % dst_ni reads (lon_min + i dlon, lat_min + jj dlat), rounds to
// the nearest source index (half away from zero), wraps the longitude index
// and clamps the latitude index.
struct RectNN {
    index_t s_ni, s_nj, d_ni;
    double s_lon0, s_lat0, s_dlon, s_dlat, d_lon0, d_lat0, d_dlon, d_dlat;
};

__global__ void nearest_rect_kernel(RectNN g, index_t n_dst, index_t *__restrict__ row_ptr, index_t *__restrict__ col, double *__restrict__ val) {
    const index_t j = static_cast<index_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (j > n_dst) return;
    row_ptr[j] = j;
    if (j == n_dst) return;
    const index_t di = j % g.d_ni, dj = j / g.d_ni;
    const double lon = g.d_lon0 + static_cast<double>(di) * g.d_dlon;
    const double lat = g.d_lat0 + static_cast<double>(dj) * g.d_dlat;
    index_t si = static_cast<index_t>(round((lon - g.s_lon0) / g.s_dlon));
    si %= g.s_ni;
    if (si < 0) si += g.s_ni;
    index_t sj = static_cast<index_t>(round((lat - g.s_lat0) / g.s_dlat));
    if (sj < 0) sj = 0;
    if (sj >= g.s_nj) sj = g.s_nj - 1;
    col[j] = sj * g.s_ni + si;
    val[j] = 1.0;
}

// get_cell_areas (weight_generator.cpp:264-313): the mesh's areas, else the
// shoelace area of the raw coordinates.
__global__ void shoelace_kernel(CellGeom g, index_t n, const double *__restrict__ given, double *__restrict__ out) {
    const index_t c = static_cast<index_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (c >= n) return;
    if (given) {
        out[c] = given[c];
        return;
    }
    const index_t b = g.off[c], m = g.off[c + 1] - b;
    if (m < 3) {
        out[c] = 0.0;
        return;
    }
    const double *lon = g.coords, *lat = g.coords + g.n_nodes;
    double s = 0.0;
    for (index_t i = 0; i < m; ++i) {
        const index_t p = g.idx[b + i], q = g.idx[b + (i + 1 == m ? 0 : i + 1)];
        s += lon[p] * lat[q] - lon[q] * lat[p];
    }
    out[c] = fabs(s) * 0.5;
}

__global__ void nn_csr_kernel(const std::int32_t *__restrict__ nearest, index_t n_dst, index_t *__restrict__ cnt) {
    const index_t j = static_cast<index_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (j < n_dst) cnt[j] = nearest[j] >= 0 ? 1 : 0;
}

__global__ void nn_fill_kernel(const std::int32_t *__restrict__ nearest, const index_t *__restrict__ row_ptr, index_t n_dst, index_t *__restrict__ col,
                               double *__restrict__ val) {
    const index_t j = static_cast<index_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (j >= n_dst || nearest[j] < 0) return;
    col[row_ptr[j]] = nearest[j];
    val[row_ptr[j]] = 1.0;
}

__global__ void fill_f64_kernel(double *__restrict__ p, index_t n, double v) {
    const index_t i = static_cast<index_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// ------------------------------------------------------------------ host
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

// The reference's float point of a cell with host libm (2D: the centroid;
// 3D: its unit vector, generate_nearest_impl<3>, :2148-2160).
template <int D>
typename PT<D>::P host_point(const Mesh &m, index_t c, SinCosMemo &memo) {
    double cx, cy;
    centroid_lonlat(m, c, memo, cx, cy);
    if constexpr (D == 2) {
        return make_float2(static_cast<float>(cx), static_cast<float>(cy));
    } else {
        const double pi = 3.14159265358979323846;
        if (m.coord_system() == CoordinateSystem::SphericalDeg) {
            cx = cx * pi / 180.0;
            cy = cy * pi / 180.0;
        }
        const geo::V3 p = memo.xyz(cx, cy);
        return make_float4(static_cast<float>(p.x), static_cast<float>(p.y), static_cast<float>(p.z), 0.0f);
    }
}

CellGeom geom_of(const Mesh &m) {
    const int cs = m.coord_system() == CoordinateSystem::SphericalDeg ? 0 : (m.coord_system() == CoordinateSystem::SphericalRad ? 1 : 2);
    return CellGeom{m.d_coords.get(), m.n_nodes(), m.d_offsets.get(), m.d_indices.get(), cs};
}

// Recomputes the listed cells' points with host libm and writes them over
// the device results.  count > kFixCap (never expected): every entry of out
// is recomputed from ids (the cell of each position; null: position = cell).
template <int D>
void fix_points(const Mesh &m, const std::int32_t *d_entry, int count, typename PT<D>::P *out, const std::int32_t *ids, index_t n_out, hipStream_t s) {
    using P = typename PT<D>::P;
    if (count <= 0) return;
    if (count <= kFixCap) {
        std::vector<std::int32_t> e(2 * static_cast<std::size_t>(count));
        AXIS_HIP_CHECK(hipMemcpyAsync(e.data(), d_entry, e.size() * sizeof(std::int32_t), hipMemcpyDeviceToHost, s));
        AXIS_HIP_CHECK(hipStreamSynchronize(s));
        std::vector<std::int32_t> pos(count);
        std::vector<P> v(count);
        SinCosMemo memo;
        for (int k = 0; k < count; ++k) {
            pos[k] = e[2 * k];
            v[k] = host_point<D>(m, e[2 * k + 1], memo);
        }
        DeviceBuffer<std::int32_t> d_pos(count);
        DeviceBuffer<P> d_v(count);
        d_pos.upload(pos.data(), count, s);
        d_v.upload(v.data(), count, s);
        scatter_fix_kernel<P><<<grid_for(count, 256), 256, 0, s>>>(d_pos.get(), d_v.get(), count, out);
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
    std::vector<P> v(n_out);
    parallel_range(n_out, [&](index_t b, index_t e) {
        SinCosMemo memo;
        for (index_t q = b; q < e; ++q) v[q] = host_point<D>(m, id[q], memo);
    });
    AXIS_HIP_CHECK(hipMemcpyAsync(out, v.data(), n_out * sizeof(P), hipMemcpyHostToDevice, s));
    AXIS_HIP_CHECK(hipStreamSynchronize(s));
}

// Points of every cell of a mesh (device), equal to the reference's host
// floats; returns how many cells were recomputed on the host.
template <int D>
index_t all_cell_points(const Mesh &m, typename PT<D>::P *out, hipStream_t s) {
    const index_t n = m.n_cells();
    if (n == 0) return 0;
    DeviceBuffer<std::int32_t> fix_entry(2 * kFixCap), fix_count(1);
    fill_int_kernel<<<1, 64, 0, s>>>(fix_count.get(), 1, 0);
    AXIS_LAUNCH_CHECK();
    centroid_select_kernel<2, D><<<grid_for(n, 256), 256, 0, s>>>(geom_of(m), n, nullptr, nullptr, out, nullptr, FixList{fix_entry.get(), fix_count.get()});
    AXIS_LAUNCH_CHECK();
    int count = 0;
    AXIS_HIP_CHECK(hipMemcpyAsync(&count, fix_count.get(), sizeof(int), hipMemcpyDeviceToHost, s));
    AXIS_HIP_CHECK(hipStreamSynchronize(s));
    fix_points<D>(m, fix_entry.get(), count, out, nullptr, n, s);
    return count;
}

// Exact nearest point for every query: pts[n] (ids[q]: the id reported for
// point q; null: q itself), queries[n_q]; out[qrow ? qrow[t] : t] = id of the
// nearest point (smallest float squared distance, then smallest id).
// Synchronises the stream before returning (its work buffers are released).
template <int D>
void nearest_points(const typename PT<D>::P *pts, const std::int32_t *ids, index_t n, const typename PT<D>::P *queries, const std::int32_t *qrow,
                    index_t n_q, std::int32_t *out, hipStream_t s) {
    using T = PT<D>;
    using P = typename T::P;
    using B = typename T::B;
    if (n == 0 || n_q == 0) return;
    const index_t n_leaves = (n + kLeafPts - 1) / kLeafPts, n_pad = n_leaves * kLeafPts;
    index_t L = 1;
    while (L < n_leaves) L *= 2;
    int height = 0;
    while ((index_t(1) << height) < L) ++height;
    if (n >= (index_t(1) << 31) || height + 1 > kStack) throw InvalidArgument("nearest search: too many points");
    DeviceBuffer<B> part(kBoundBlocks), bounds(1);
    bounds_kernel<D><<<kBoundBlocks, 256, 0, s>>>(pts, n, part.get());
    AXIS_LAUNCH_CHECK();
    bounds_fold_kernel<D><<<1, kBoundBlocks, 0, s>>>(part.get(), bounds.get());
    AXIS_LAUNCH_CHECK();
    DeviceBuffer<std::uint32_t> key(n), kpos(n);
    point_keys_kernel<D><<<grid_for(n, 256), 256, 0, s>>>(pts, n, bounds.get(), key.get(), kpos.get());
    AXIS_LAUNCH_CHECK();
    radix_sort_pairs_u32(key.get(), kpos.get(), static_cast<std::size_t>(n), D * T::kKeyBits, s);
    DeviceBuffer<P> spts(n_pad);
    DeviceBuffer<std::int32_t> sidx(n_pad);
    gather_points_kernel<D><<<grid_for(n_pad, 256), 256, 0, s>>>(kpos.get(), pts, ids, n, n_pad, spts.get(), sidx.get());
    AXIS_LAUNCH_CHECK();
    DeviceBuffer<B> box(2 * L);
    leaf_boxes_kernel<D><<<grid_for(L, 256), 256, 0, s>>>(spts.get(), n, n_leaves, L, box.get());
    AXIS_LAUNCH_CHECK();
    for (index_t lo = L / 2; lo >= 1; lo /= 2) {
        inner_boxes_kernel<D><<<grid_for(lo, 256), 256, 0, s>>>(box.get(), lo);
        AXIS_LAUNCH_CHECK();
    }
    // queries in their own order: destination and dry-row indices are
    // already spatially coherent (a Morton sort of them measured slower)
    nearest_kernel<D><<<grid_for(n_q, kNearThreads), kNearThreads, 0, s>>>(box.get(), spts.get(), sidx.get(), n_leaves, L, queries, qrow, nullptr,
                                                                           n_q, out);
    AXIS_LAUNCH_CHECK();
    AXIS_HIP_CHECK(hipStreamSynchronize(s));
}

}  // namespace

index_t cell_centroids_f32(const Mesh &m, float2 *out, hipStream_t s) { return all_cell_points<2>(m, out, s); }

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
        centroid_select_kernel<0, 2><<<grid_for(n_src, 256), 256, 0, s>>>(geom_of(src), n_src, src.d_mask.get(), pos_s.get(), p0.get(), pid0.get(),
                                                                          FixList{fix_entry.get(), ints.get() + 1});
        AXIS_LAUNCH_CHECK();
        centroid_select_kernel<1, 2><<<grid_for(n_dst, 256), 256, 0, s>>>(geom_of(dst), n_dst, wet_cnt.get(), pos_d.get(), q.get(), qrow.get(),
                                                                          FixList{fix_entry.get() + 2 * kFixCap, ints.get() + 2});
        AXIS_LAUNCH_CHECK();
        index_t n_wet = 0, n_q = 0;
        AXIS_HIP_CHECK(hipMemcpyAsync(&n_wet, pos_s.get() + n_src, sizeof(index_t), hipMemcpyDeviceToHost, s));
        AXIS_HIP_CHECK(hipMemcpyAsync(&n_q, pos_d.get() + n_dst, sizeof(index_t), hipMemcpyDeviceToHost, s));
        AXIS_HIP_CHECK(hipMemcpyAsync(h_ints, ints.get(), sizeof h_ints, hipMemcpyDeviceToHost, s));
        AXIS_HIP_CHECK(hipStreamSynchronize(s));
        tr.mark("wet_rows+centroids");
        if (n_wet > 0 && n_q > 0) {
            fix_points<2>(src, fix_entry.get(), h_ints[1], p0.get(), pid0.get(), n_wet, s);
            fix_points<2>(dst, fix_entry.get() + 2 * kFixCap, h_ints[2], q.get(), qrow.get(), n_q, s);
            tr.mark("host_fixups");
            nearest_points<2>(p0.get(), pid0.get(), n_wet, q.get(), qrow.get(), n_q, nearest.get(), s);
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
    M.reset_apply_layout();
    M.len_hist_ready = false;
}

std::unique_ptr<InterpolationMatrix> generate_nearest(const Mesh &src, const Mesh &dst, const RegridConfig &cfg, const MeshGeom *src_reg,
                                                      const MeshGeom *dst_reg, hipStream_t s) {
    const auto t0 = std::chrono::steady_clock::now();
    const index_t n_src = src.n_cells(), n_dst = dst.n_cells();
    auto M = std::make_unique<InterpolationMatrix>();
    M->n_src = n_src;
    M->n_dst = n_dst;
    const index_t ns = std::max<index_t>(n_src, 1), nd = std::max<index_t>(n_dst, 1);
    M->frac_a.resize(ns);
    M->frac_b.resize(nd);
    M->area_a.resize(ns);
    M->area_b.resize(nd);
    M->row_ptr.resize(n_dst + 1);
    if (src_reg && dst_reg) {
        // generate_nearest_rect: fractions and areas stay zero (:2070-2080)
        RectNN g{src_reg->ni,        src_reg->nj,        dst_reg->ni,        src_reg->lon_min,   src_reg->lat_min,   src_reg->delta_lon,
                 src_reg->delta_lat, dst_reg->lon_min,   dst_reg->lat_min,   dst_reg->delta_lon, dst_reg->delta_lat};
        M->nnz = n_dst;
        M->col.resize(nd);
        M->val.resize(nd);
        nearest_rect_kernel<<<grid_for(n_dst + 1, 256), 256, 0, s>>>(g, n_dst, M->row_ptr.get(), M->col.get(), M->val.get());
        AXIS_LAUNCH_CHECK();
        fill_f64_kernel<<<grid_for(ns, 256), 256, 0, s>>>(M->frac_a.get(), ns, 0.0);
        fill_f64_kernel<<<grid_for(nd, 256), 256, 0, s>>>(M->frac_b.get(), nd, 0.0);
        fill_f64_kernel<<<grid_for(ns, 256), 256, 0, s>>>(M->area_a.get(), ns, 0.0);
        fill_f64_kernel<<<grid_for(nd, 256), 256, 0, s>>>(M->area_b.get(), nd, 0.0);
        AXIS_LAUNCH_CHECK();
        M->max_row_nnz = n_dst > 0 ? 1 : 0;
    } else {
        // generate_nearest_impl<3> on a sphere, <2> in the plane (:2129-2271)
        const bool sph = src.coord_system() != CoordinateSystem::Cartesian3D;
        DeviceBuffer<std::int32_t> nearest(nd);
        fill_int_kernel<<<grid_for(nd, 256), 256, 0, s>>>(nearest.get(), nd, -1);
        AXIS_LAUNCH_CHECK();
        if (sph) {
            DeviceBuffer<float4> sp(ns), dp(nd);
            all_cell_points<3>(src, sp.get(), s);
            all_cell_points<3>(dst, dp.get(), s);
            nearest_points<3>(sp.get(), nullptr, n_src, dp.get(), nullptr, n_dst, nearest.get(), s);
        } else {
            DeviceBuffer<float2> sp(ns), dp(nd);
            all_cell_points<2>(src, sp.get(), s);
            all_cell_points<2>(dst, dp.get(), s);
            nearest_points<2>(sp.get(), nullptr, n_src, dp.get(), nullptr, n_dst, nearest.get(), s);
        }
        DeviceBuffer<index_t> cnt(nd + 1);
        nn_csr_kernel<<<grid_for(nd, 256), 256, 0, s>>>(nearest.get(), n_dst, cnt.get());
        AXIS_LAUNCH_CHECK();
        exclusive_scan_i64(cnt.get(), M->row_ptr.get(), n_dst, s);
        index_t nnz = 0;
        AXIS_HIP_CHECK(hipMemcpyAsync(&nnz, M->row_ptr.get() + n_dst, sizeof(index_t), hipMemcpyDeviceToHost, s));
        AXIS_HIP_CHECK(hipStreamSynchronize(s));
        if (nnz < n_dst && cfg.unmapped == UnmappedAction::Error)
            throw InvalidArgument("WeightGenerator::generate_nearest: unmapped destination cell 0");
        M->nnz = nnz;
        M->col.resize(std::max<index_t>(nnz, 1));
        M->val.resize(std::max<index_t>(nnz, 1));
        nn_fill_kernel<<<grid_for(nd, 256), 256, 0, s>>>(nearest.get(), M->row_ptr.get(), n_dst, M->col.get(), M->val.get());
        AXIS_LAUNCH_CHECK();
        fill_f64_kernel<<<grid_for(ns, 256), 256, 0, s>>>(M->frac_a.get(), ns, 1.0);
        fill_f64_kernel<<<grid_for(nd, 256), 256, 0, s>>>(M->frac_b.get(), nd, 1.0);
        AXIS_LAUNCH_CHECK();
        shoelace_kernel<<<grid_for(n_src, 256), 256, 0, s>>>(geom_of(src), n_src, src.has_areas() ? src.d_areas.get() : nullptr, M->area_a.get());
        shoelace_kernel<<<grid_for(n_dst, 256), 256, 0, s>>>(geom_of(dst), n_dst, dst.has_areas() ? dst.d_areas.get() : nullptr, M->area_b.get());
        AXIS_LAUNCH_CHECK();
        M->max_row_nnz = nnz > 0 ? 1 : 0;
        AXIS_HIP_CHECK(hipStreamSynchronize(s));  // nearest is released on scope exit
    }
    cell_morton_keys(dst, M->row_key, s);
    AXIS_HIP_CHECK(hipStreamSynchronize(s));
    M->gen_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return M;
}

}  // namespace axis_hip::detail
