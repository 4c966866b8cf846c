// gridquery.hip -- broad phase for a rectilinear degree source: the source
// cells' (lon, lat) index grid replaces the BVH.
//
// Result: exactly lbvh_query's -- for every destination cell the source cells
// whose float boxes intersect its box (closed intervals, as ArborX::intersects
// in the reference's BVH query, weight_generator.cpp:3345-3367) -- in another
// order within a row (the CSR assembly sorts each row by column).
//
// Why the enumeration is complete.  Let a source box B_s (cell (i, j): lon
// values ulon[i], ulon[i+1]) and a destination box B_d share a point p.
//  * z: p.z lies in both boxes' z-ranges, so row j's band [min z-lo, max z-hi]
//    over its cells meets B_d's z-range.  Rows come from a binary search over
//    the suffix minima of the bands' lower ends and the prefix maxima of their
//    upper ends (both nondecreasing), a superset of those rows.
//  * longitude: unless one box's xy-rectangle comes within 1e-6 of the polar
//    axis (then every longitude is taken), p's angle lies in the angular spans
//    of both rectangles; B_s's span lies within [ulon[i] - dlt_j, ulon[i+1] +
//    dlt_j], where dlt_j is the largest overhang of a row-j box's span beyond
//    its cell's longitudes (a box whose rectangle nears the axis makes the
//    whole row taken).  So i runs over the cells whose widened longitude
//    interval meets B_d's span (mod 360), widened by 1e-7 degrees for the
//    rounding of atan2.
//  * distance from the polar axis: |p.xy| lies in both rectangles' distance
//    ranges, so row j's radial band (the union of its boxes' ranges, rounded
//    outward to float) meets B_d's range (widened by 1e-9); near the poles,
//    where the boxes' z-bands span many rows, this keeps the rows at B_d's
//    distance from the axis.
// Every enumerated cell then takes the same float box test the BVH applies.
//
//   1. node_class_kernel: each node's axis classes (largest tolerance-unique
//      value not above the coordinate), the axes in LDS;
//   2. cell_grid_kernel: each cell's lower-left (i, j), checked to be a 2 x 2
//      node block; cell_at[j (nlon-1) + i] = cell; per-row z band, longitude
//      overhang and radial band (order-preserving integer keys, atomics
//      after a wave reduction);
//   3. row_finish_kernel: the keys to floats, suffix minima / prefix maxima;
//   4. grid_count_kernel: one lane per destination cell counts its hits
//      (the row tables and, when it fits, the longitude axis in LDS); cells
//      with more than kLaneItems grid cells to test are listed for
//      grid_long_kernel (one wave per cell: lanes over its rows, then over
//      its rows' grid cells taken together), which passes cells of more than
//      64 rows or kHeavyItems grid cells on to grid_heavy_kernel (one
//      1024-thread workgroup per cell); then an exclusive scan of the counts,
//      and the same kernels fill the candidate list and each candidate's
//      destination row, every cell's hits in row, range and column order.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>

#include "axis_hip.hpp"
#include "host_util.hpp"

namespace axis_hip::detail {

namespace {

constexpr int kLaneItems = 64;        // grid cells one lane tests before the cell goes to a wave
constexpr int kUnroll = 4;            // box loads a lane keeps in flight
constexpr int kMaxLonRanges = 3;      // longitude index ranges per row (wraps of the span)
constexpr double kAxisGap = 1e-6;     // rectangle this near the polar axis: every longitude
constexpr double kLonMarginDeg = 1e-7;     // source overhangs (double arithmetic)
constexpr float kDstLonMarginDeg = 1e-3f;  // destination spans (single precision)
constexpr float kAllLon = 1e30f;      // row overhang meaning "every cell of the row"
constexpr double kRad2Deg = 180.0 / kPi;
constexpr int kQueryBlock = 256;
constexpr int kLongBlocks = 512;      // workgroups of grid_long_kernel (they loop over the list)
constexpr index_t kCountBlocks = 2048;  // workgroups of grid_count_kernel (they loop over the cells)
constexpr std::size_t kMaxLds = 64 * 1024;
constexpr std::size_t kBigTables = 32 * 1024;  // LDS tables beyond this: the counting pass takes 1024-thread workgroups
constexpr int kFillWin = 512;      // candidate-list window per wave of the filling pass (entries), small tables
constexpr int kFillWinBig = 1024;  // the same beside tables of more than kBigTables (one 1024-thread workgroup per CU)
constexpr std::size_t fill_lds(int block, int win) { return static_cast<std::size_t>(block / 64) * (win + 64) * sizeof(std::int32_t); }
// per cell row: z-band suffix minimum and prefix maximum, longitude overhang,
// radial band (distance of the boxes' xy-rectangles from the polar axis)
constexpr int kRowFields = 5;

__device__ inline int lower_bound_d(const double *v, int n, double x) {  // first v[k] >= x
    int lo = 0, hi = n;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (v[mid] < x)
            lo = mid + 1;
        else
            hi = mid;
    }
    return lo;
}

__device__ inline int upper_bound_d(const double *v, int n, double x) {  // first v[k] > x
    int lo = 0, hi = n;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (v[mid] <= x)
            lo = mid + 1;
        else
            hi = mid;
    }
    return lo;
}

template <typename T>
__device__ inline int lower_bound_t(const T *v, int n, T x) {
    int lo = 0, hi = n;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (v[mid] < x)
            lo = mid + 1;
        else
            hi = mid;
    }
    return lo;
}

template <typename T>
__device__ inline int upper_bound_t(const T *v, int n, T x) {
    int lo = 0, hi = n;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (v[mid] <= x)
            lo = mid + 1;
        else
            hi = mid;
    }
    return lo;
}

// floats as unsigned keys of the same order
__device__ inline unsigned fkey(float f) {
    const unsigned b = __float_as_uint(f);
    return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}
__device__ inline float fkey_inv(unsigned k) { return __uint_as_float((k & 0x80000000u) ? (k & 0x7FFFFFFFu) : ~k); }

// angular span (degrees) of an xy-rectangle about reference angle ref_deg:
// [ref + lo, ref + hi]; false when the rectangle comes within kAxisGap of the axis
__device__ inline bool rect_span(double x0, double x1, double y0, double y1, double ref_deg, double &lo, double &hi) {
    if (x0 <= kAxisGap && x1 >= -kAxisGap && y0 <= kAxisGap && y1 >= -kAxisGap) return false;
    lo = 1e300;
    hi = -1e300;
    for (int q = 0; q < 4; ++q) {
        double d = atan2((q & 2) ? y1 : y0, (q & 1) ? x1 : x0) * kRad2Deg - ref_deg;
        d -= 360.0 * rint(d / 360.0);  // into [-180, 180]
        lo = fmin(lo, d);
        hi = fmax(hi, d);
    }
    return true;
}

// distance range [r0, r1] of an xy-rectangle's points from the polar axis
__device__ inline void rect_rho(double x0, double x1, double y0, double y1, double &r0, double &r1) {
    const double cx = x0 > 0.0 ? x0 : (x1 < 0.0 ? -x1 : 0.0), cy = y0 > 0.0 ? y0 : (y1 < 0.0 ? -y1 : 0.0);
    r0 = sqrt(cx * cx + cy * cy);
    const double fx = fmax(fabs(x0), fabs(x1)), fy = fmax(fabs(y0), fabs(y1));
    r1 = sqrt(fx * fx + fy * fy);
}

// ---- 1. node axis classes
__global__ void node_class_kernel(index_t n_nodes, const double *lon, const double *lat, const double *axes, int nlon, int nlat, std::int32_t *na,
                                  std::int32_t *nb) {
    extern __shared__ double s_axes[];
    for (int k = threadIdx.x; k < nlon + nlat; k += blockDim.x) s_axes[k] = axes[k];
    __syncthreads();
    const index_t i = static_cast<index_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n_nodes) return;
    na[i] = upper_bound_d(s_axes, nlon, lon[i]) - 1;
    nb[i] = upper_bound_d(s_axes + nlon, nlat, lat[i]) - 1;
}

// ---- 2. cells on the grid, row bands and overhangs
__global__ void __launch_bounds__(256) cell_grid_kernel(index_t n_cells, const index_t *off, const index_t *ind, const std::int32_t *na,
                                                        const std::int32_t *nb, const double *ulon, int nlon, int nlat, const float *box,
                                                        std::int32_t *cell_at, int *bad, unsigned *row_keys) {
    const index_t c = static_cast<index_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    // every lane stays to the wave reduction below; lanes without a valid
    // cell carry neutral keys
    bool ok = c < n_cells;
    int i = 0, j = -1;
    if (ok) {
        const index_t b = off[c];
        ok = off[c + 1] - b == 4;
        int ia[4] = {0, 0, 0, 0}, ib[4] = {0, 0, 0, 0};
        for (int v = 0; v < 4 && ok; ++v) {
            const index_t nd = ind[b + v];
            ia[v] = na[nd];
            ib[v] = nb[nd];
        }
        i = min(min(ia[0], ia[1]), min(ia[2], ia[3]));
        j = min(min(ib[0], ib[1]), min(ib[2], ib[3]));
        ok = ok && i >= 0 && j >= 0 && i < nlon - 1 && j < nlat - 1;
        unsigned seen = 0;  // bit (di + 2 dj) for each corner
        for (int v = 0; v < 4 && ok; ++v) {
            const int di = ia[v] - i, dj = ib[v] - j;
            ok = di <= 1 && dj <= 1;
            if (ok) seen |= 1u << (di + 2 * dj);
        }
        ok = ok && seen == 0xFu;
        const std::int64_t slot = static_cast<std::int64_t>(j) * (nlon - 1) + i;
        if (ok && atomicCAS(&cell_at[slot], -1, static_cast<std::int32_t>(c)) != -1) ok = false;
        if (ok && slot != c) atomicOr(bad, 2);  // (not in grid order)
        if (!ok) {
            atomicOr(bad, 1);
            j = -1;
        }
    }
    unsigned k0 = 0xFFFFFFFFu, k1 = 0u, k2 = 0u, k3 = 0xFFFFFFFFu, k4 = 0u;
    if (ok) {
        const float *bx = box + 6 * c;
        // overhang of the box's angular span beyond [ulon[i], ulon[i+1]]
        const double l0 = ulon[i], l1 = ulon[i + 1], mid = 0.5 * (l0 + l1);
        double lo, hi;
        float dlt = kAllLon;
        if (l1 - l0 < 90.0 && rect_span(bx[0], bx[3], bx[1], bx[4], mid, lo, hi))
            dlt = static_cast<float>(fmax(fmax((l0 - mid) - lo, hi - (l1 - mid)), 0.0) * (1.0 + 1e-6) + kLonMarginDeg);
        k0 = fkey(bx[2]);
        k1 = fkey(bx[5]);
        k2 = fkey(dlt);
        double r0, r1;
        rect_rho(bx[0], bx[3], bx[1], bx[4], r0, r1);
        k3 = fkey(__double2float_rd(r0));
        k4 = fkey(__double2float_ru(r1));
    }
    const int rows = nlat - 1;
    // a wave whose valid cells share one row (cells in grid order) reduces
    // first and updates once
    const int j0 = __shfl(j, 0);
    const bool uniform = __ballot(j != j0 && j >= 0) == 0 && j0 >= 0;
    if (uniform) {
        for (int o = 32; o > 0; o >>= 1) {
            k0 = min(k0, static_cast<unsigned>(__shfl_xor(static_cast<int>(k0), o)));
            k1 = max(k1, static_cast<unsigned>(__shfl_xor(static_cast<int>(k1), o)));
            k2 = max(k2, static_cast<unsigned>(__shfl_xor(static_cast<int>(k2), o)));
            k3 = min(k3, static_cast<unsigned>(__shfl_xor(static_cast<int>(k3), o)));
            k4 = max(k4, static_cast<unsigned>(__shfl_xor(static_cast<int>(k4), o)));
        }
        if ((threadIdx.x & 63) != 0) return;
        j = j0;
    } else if (j < 0) {
        return;
    }
    atomicMin(&row_keys[j], k0);
    atomicMax(&row_keys[rows + j], k1);
    atomicMax(&row_keys[2 * rows + j], k2);
    atomicMin(&row_keys[3 * rows + j], k3);
    atomicMax(&row_keys[4 * rows + j], k4);
}

// ---- 3. row tables: zlo as suffix minima, zhi as prefix maxima, overhang
// (one wave: each lane scans a chunk, the lanes combine by shuffles)
__global__ void __launch_bounds__(64) row_finish_kernel(int rows, const unsigned *row_keys, float *row_tab) {
    const int lane = threadIdx.x;
    const int per = (rows + 63) / 64;
    const int b = min(lane * per, rows), e = min(b + per, rows);
    float mn = INFINITY, mx = -INFINITY;
    for (int j = b; j < e; ++j) {
        mn = fminf(mn, fkey_inv(row_keys[j]));
        mx = fmaxf(mx, fkey_inv(row_keys[rows + j]));
        for (int f = 2; f < kRowFields; ++f) row_tab[f * rows + j] = fkey_inv(row_keys[f * rows + j]);
    }
    // suffix minimum over the chunks after this lane's, prefix maximum before it
    float suf = mn, pre = mx;
    for (int o = 1; o < 64; o <<= 1) {
        const float a = __shfl_down(suf, o), c = __shfl_up(pre, o);
        if (lane + o < 64) suf = fminf(suf, a);
        if (lane >= o) pre = fmaxf(pre, c);
    }
    float after = __shfl_down(suf, 1), before = __shfl_up(pre, 1);
    if (lane == 63) after = INFINITY;
    if (lane == 0) before = -INFINITY;
    for (int j = e - 1; j >= b; --j) {
        after = fminf(after, fkey_inv(row_keys[j]));
        row_tab[j] = after;
    }
    for (int j = b; j < e; ++j) {
        before = fmaxf(before, fkey_inv(row_keys[rows + j]));
        row_tab[rows + j] = before;
    }
}

// ---- 4. queries
struct GridView {
    const double *ulon;  // tolerance-unique longitudes, ascending (degrees)
    const float *rows;   // kRowFields arrays of nrow floats
    int nlon, nrow;      // lon node count; cell rows
    bool lon_lds;        // ulon copied to LDS (else read from HBM)
    double u0, inv_du;   // ulon[0] and the inverse mean spacing (index guesses)
    const std::int32_t *cell_at;  // null: cell j (nlon - 1) + i is grid slot (i, j)
    const float *src_box;
};

// tables in LDS: the row arrays (floats), then ulon (doubles) when it fits
struct Tables {
    const float *zlo, *zhi, *dlt, *rlo, *rhi;
    const double *ulon;
};
__device__ inline Tables load_tables(const GridView &G, double *s) {
    float *f = reinterpret_cast<float *>(s);
    for (int k = threadIdx.x; k < kRowFields * G.nrow; k += blockDim.x) f[k] = G.rows[k];
    double *u = s + (kRowFields * G.nrow + 1) / 2;
    if (G.lon_lds)
        for (int k = threadIdx.x; k < G.nlon; k += blockDim.x) u[k] = G.ulon[k];
    __syncthreads();
    return Tables{f, f + G.nrow, f + 2 * G.nrow, f + 3 * G.nrow, f + 4 * G.nrow, G.lon_lds ? u : G.ulon};
}

struct Dst {
    float bx[6];
    int j0, j1;    // candidate rows
    bool all_lon;  // rectangle near the axis
    float a0, a1;  // angular span (degrees, widened)
    float r0, r1;  // distance range from the axis (widened)
};

__device__ inline void dst_setup(const GridView &G, const Tables &T, const float *box, index_t d, Dst &D) {
    for (int a = 0; a < 6; ++a) D.bx[a] = box[6 * d + a];
    D.j0 = 0;
    D.j1 = -1;
    if (!(D.bx[0] <= D.bx[3] && D.bx[1] <= D.bx[4] && D.bx[2] <= D.bx[5])) return;  // empty (or NaN) box: no hit
    D.j0 = lower_bound_t(T.zhi, G.nrow, D.bx[2]);     // first row whose band top reaches the box
    D.j1 = upper_bound_t(T.zlo, G.nrow, D.bx[5]) - 1;  // last row whose band bottom is below the box top
    // single precision with margins far above its rounding (a wider range
    // only adds grid cells to test)
    const float *b = D.bx;
    const float cx = b[0] > 0.0f ? b[0] : (b[3] < 0.0f ? -b[3] : 0.0f), cy = b[1] > 0.0f ? b[1] : (b[4] < 0.0f ? -b[4] : 0.0f);
    const float fx = fmaxf(fabsf(b[0]), fabsf(b[3])), fy = fmaxf(fabsf(b[1]), fabsf(b[4]));
    D.r0 = sqrtf(cx * cx + cy * cy) * (1.0f - 1e-5f) - 1e-7f;
    D.r1 = sqrtf(fx * fx + fy * fy) * (1.0f + 1e-5f) + 1e-7f;
    D.all_lon = b[0] <= static_cast<float>(kAxisGap) && b[3] >= -static_cast<float>(kAxisGap) && b[1] <= static_cast<float>(kAxisGap) &&
                b[4] >= -static_cast<float>(kAxisGap);
    if (D.all_lon) return;
    const float rad2deg = static_cast<float>(kRad2Deg);
    const float ref = atan2f(0.5f * (b[1] + b[4]), 0.5f * (b[0] + b[3])) * rad2deg;
    float lo = 0.0f, hi = 0.0f;
    for (int q = 0; q < 4; ++q) {
        float d = atan2f((q & 2) ? b[4] : b[1], (q & 1) ? b[3] : b[0]) * rad2deg - ref;
        d -= 360.0f * rintf(d * (1.0f / 360.0f));  // into [-180, 180]
        lo = fminf(lo, d);
        hi = fmaxf(hi, d);
    }
    D.a0 = ref + lo - kDstLonMarginDeg;
    D.a1 = ref + hi + kDstLonMarginDeg;
}

// first i in [0, last] with ulon[i + 1] >= w (last + 1 if none): a guess from
// the mean spacing walked a few steps, else a binary search
__device__ inline int lon_first(const GridView &G, const double *ulon, double w) {
    const int last = G.nlon - 2;
    int g = static_cast<int>(fmin(fmax(floor((w - G.u0) * G.inv_du), 0.0), static_cast<double>(last + 1)));
    for (int step = 0; step < 8; ++step) {
        if (g > 0 && ulon[g] >= w)
            --g;
        else if (g <= last && ulon[g + 1] < w)
            ++g;
        else
            return g;
    }
    return max(lower_bound_d(ulon, G.nlon, w) - 1, 0);
}
// last i in [0, last] with ulon[i] <= w (-1 if none)
__device__ inline int lon_last(const GridView &G, const double *ulon, double w) {
    const int last = G.nlon - 2;
    int g = static_cast<int>(fmin(fmax(floor((w - G.u0) * G.inv_du), -1.0), static_cast<double>(last)));
    for (int step = 0; step < 8; ++step) {
        if (g < last && ulon[g + 1] <= w)
            ++g;
        else if (g >= 0 && ulon[g] > w)
            --g;
        else
            return g;
    }
    return min(upper_bound_d(ulon, G.nlon, w) - 1, last);
}

// longitude cell ranges of row j for destination D; returns their number (0:
// the row's radial band misses the box's)
__device__ inline int row_ranges(const GridView &G, const Tables &T, int j, const Dst &D, int *r0, int *r1) {
    if (T.rlo[j] > D.r1 || T.rhi[j] < D.r0) return 0;
    const double *ulon = T.ulon;
    const float dlt = T.dlt[j];
    const int last = G.nlon - 2;
    const double w0 = D.a0 - dlt, w1 = D.a1 + dlt;
    // every cell: near the axis, or a span so wide that a cell could meet two wraps
    if (D.all_lon || dlt >= 90.0f || w1 - w0 >= 180.0) {
        r0[0] = 0;
        r1[0] = last;
        return 1;
    }
    const double kmin = ceil((ulon[0] - w1) * (1.0 / 360.0) - 1e-12), kmax = floor((ulon[G.nlon - 1] - w0) * (1.0 / 360.0) + 1e-12);
    if (kmax - kmin + 1 > kMaxLonRanges) {
        r0[0] = 0;
        r1[0] = last;
        return 1;
    }
    int n = 0;
    for (double k = kmin; k <= kmax; k += 1.0) {
        // cells whose [ulon[i], ulon[i+1]] meets [w0, w1] + 360 k
        const int i0 = lon_first(G, ulon, w0 + 360.0 * k), i1 = lon_last(G, ulon, w1 + 360.0 * k);
        if (i0 <= i1) {
            r0[n] = i0;
            r1[n] = i1;
            ++n;
        }
    }
    return n;
}

__device__ inline bool boxes_meet(const float *a, const float *b) {
    return !(a[3] < b[0] || b[3] < a[0] || a[4] < b[1] || b[4] < a[1] || a[5] < b[2] || b[5] < a[2]);
}

__device__ inline std::int32_t cell_of(const GridView &G, std::int64_t k) { return G.cell_at ? G.cell_at[k] : static_cast<std::int32_t>(k); }

// orders the LDS accesses of a wave's lanes around it (a wave's LDS
// operations complete in program order; this keeps the compiler's order)
__device__ inline void wave_lds_fence() { asm volatile("" ::: "memory"); }

// kFill: every hit goes to put(position, source cell), positions from
// cand_off[d] on
template <bool kFill, class Put>
__device__ inline void count_one(const GridView &G, const Tables &T, const float *dst_box, index_t d, index_t *count, const index_t *cand_off,
                                 std::int32_t *long_rows, int *n_long, Put &&put) {
    Dst D;
    dst_setup(G, T, dst_box, d, D);
    int r0[kMaxLonRanges], r1[kMaxLonRanges];
    // items first: a cell with many goes to a wave
    int items = 0;
    for (int j = D.j0; j <= D.j1 && items <= kLaneItems; ++j) {
        const int n = row_ranges(G, T, j, D, r0, r1);
        for (int q = 0; q < n; ++q) items += r1[q] - r0[q] + 1;
    }
    if (items > kLaneItems) {
        if (!kFill) {
            count[d] = 0;
            long_rows[atomicAdd(n_long, 1)] = static_cast<std::int32_t>(d);
        }
        return;
    }
    index_t pos = kFill ? cand_off[d] : 0;
    for (int j = D.j0; j <= D.j1; ++j) {
        const int n = row_ranges(G, T, j, D, r0, r1);
        const std::int64_t base = static_cast<std::int64_t>(j) * (G.nlon - 1);
        for (int q = 0; q < n; ++q)
            for (int a = r0[q]; a <= r1[q]; a += kUnroll) {
                // kUnroll boxes in flight, then their tests
                float bs[kUnroll][6];
                std::int32_t cs[kUnroll];
#pragma unroll
                for (int u = 0; u < kUnroll; ++u) {
                    cs[u] = a + u <= r1[q] ? cell_of(G, base + a + u) : -1;
                    if (cs[u] >= 0) {
                        const float *p = G.src_box + 6 * static_cast<std::int64_t>(cs[u]);
#pragma unroll
                        for (int t = 0; t < 6; ++t) bs[u][t] = p[t];
                    }
                }
#pragma unroll
                for (int u = 0; u < kUnroll; ++u)
                    if (cs[u] >= 0 && boxes_meet(bs[u], D.bx)) {
                        if (kFill) put(pos, cs[u]);
                        ++pos;
                    }
            }
    }
    if (!kFill) count[d] = pos;
}

// counting: one lane per destination cell; workgroups loop over the cells
// (the tables are loaded once per workgroup)
template <int BLOCK>
__global__ void __launch_bounds__(BLOCK) grid_count_kernel(GridView G, const float *dst_box, index_t n_dst, index_t *count,
                                                                 std::int32_t *long_rows, int *n_long) {
    extern __shared__ double s_tab[];
    const Tables T = load_tables(G, s_tab);
    for (index_t d = static_cast<index_t>(blockIdx.x) * BLOCK + threadIdx.x; d < n_dst; d += static_cast<index_t>(gridDim.x) * BLOCK)
        count_one<false>(G, T, dst_box, d, count, nullptr, long_rows, n_long, [](index_t, std::int32_t) {});
}

// filling: one lane per destination cell as in the count, the hits of a
// wave's 64 cells staged in a window of W entries in LDS (the cells' ranges
// of the candidate list are consecutive, so position - the first cell's
// offset indexes the window), then written out in position order by the
// wave's lanes together (coalesced), each entry's row found by a binary
// search over the 64 offsets.  Hits past the window are stored directly.
// Entries of the listed cells (their hits come from grid_long_kernel) stay
// marked -1 in the window and are not written.
template <int BLOCK, int W>
__global__ void __launch_bounds__(BLOCK) grid_fill_kernel(GridView G, const float *dst_box, index_t n_dst, const index_t *cand_off,
                                                                std::int32_t *cand_src, std::int32_t *pair_row, unsigned tab_bytes) {
    extern __shared__ double s_tab[];
    const Tables T = load_tables(G, s_tab);
    const int lane = static_cast<int>(threadIdx.x & 63), wv = static_cast<int>(threadIdx.x >> 6);
    std::int32_t *win = reinterpret_cast<std::int32_t *>(reinterpret_cast<char *>(s_tab) + tab_bytes) + wv * (W + 64);
    std::int32_t *offs = win + W;
    for (index_t d0 = static_cast<index_t>(blockIdx.x) * BLOCK + wv * 64; d0 < n_dst; d0 += static_cast<index_t>(gridDim.x) * BLOCK) {
        const index_t d = d0 + lane;
        const bool in = d < n_dst;
        const index_t base = cand_off[d0];
        const int span = static_cast<int>(min<index_t>(cand_off[min<index_t>(d0 + 64, n_dst)] - base, W));
        for (int i = lane; i < span; i += 64) win[i] = -1;
        offs[lane] = in ? static_cast<std::int32_t>(min<index_t>(cand_off[d] - base, index_t(0x7fffffff))) : 0x7fffffff;
        wave_lds_fence();
        if (in)
            count_one<true>(G, T, dst_box, d, nullptr, cand_off, nullptr, nullptr, [&](index_t pos, std::int32_t c) {
                const index_t k = pos - base;
                if (k < W) {
                    win[k] = c;
                } else {
                    cand_src[pos] = c;
                    pair_row[pos] = static_cast<std::int32_t>(d);
                }
            });
        wave_lds_fence();
        for (int i = lane; i < span; i += 64) {
            const std::int32_t c = win[i];
            if (c < 0) continue;
            int r = 0;  // the last lane whose offset is at most i
#pragma unroll
            for (int step = 32; step >= 1; step >>= 1)
                if (offs[r + step] <= i) r += step;
            cand_src[base + i] = c;
            pair_row[base + i] = static_cast<std::int32_t>(d0 + r);
        }
        wave_lds_fence();
    }
}

// One wave per listed destination cell (more than kLaneItems grid cells to
// test; near the source's poles: at 0.1 deg -> C768, 121,632 cells of ~10
// rows and ~115 grid cells each).  The lanes first take one row each (its
// longitude ranges and grid-cell count, kept in LDS, and a prefix scan of the
// counts), then one grid cell each of the cell's rows taken together, found
// by a binary search over the row offsets; the hits come out in row, range
// and column order, written in that order (ballot).  Workgroups loop over
// the list.
constexpr int kLongBlock = 1024;  // (16 waves: per-wave LDS beside the tables, two workgroups per CU)
constexpr int kLongWaves = kLongBlock / 64;
constexpr int kHeavyItems = 512;  // more grid cells than this: a workgroup per cell (grid_heavy_kernel)

// a compiler barrier: LDS written by some lanes of a wave and read by others
// after it, the wave's LDS instructions being executed in program order

template <bool kFill>
__global__ void __launch_bounds__(kLongBlock) grid_long_kernel(GridView G, const float *dst_box, const std::int32_t *long_rows, const int *n_long,
                                                                index_t *count, const index_t *cand_off, std::int32_t *cand_src,
                                                                std::int32_t *pair_row, std::int32_t *heavy_rows, int *n_heavy) {
    extern __shared__ double s_tab[];
    __shared__ int s_off[kLongWaves][65];                    // grid cells before row lane L
    __shared__ int s_lo[kLongWaves][kMaxLonRanges][64], s_hi[kLongWaves][kMaxLonRanges][64];
    __shared__ int s_nq[kLongWaves][64];
    const Tables T = load_tables(G, s_tab);
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int nl = *n_long;
    for (int li = blockIdx.x * kLongWaves + wv; li < nl; li += gridDim.x * kLongWaves) {
        const index_t d = long_rows[li];
        Dst D;
        dst_setup(G, T, dst_box, d, D);
        index_t pos = kFill ? cand_off[d] : 0;
        for (int jb = D.j0; jb <= D.j1; jb += 64) {
            // one row per lane
            int r0[kMaxLonRanges], r1[kMaxLonRanges];
            int nq = 0, cnt = 0;
            if (jb + lane <= D.j1) {
                nq = row_ranges(G, T, jb + lane, D, r0, r1);
                for (int q = 0; q < nq; ++q) cnt += r1[q] - r0[q] + 1;
            }
            int inc = cnt;
            for (int o = 1; o < 64; o <<= 1) {
                const int t = __shfl_up(inc, o, 64);
                if (lane >= o) inc += t;
            }
            const int total = __shfl(inc, 63, 64);
            if (jb == D.j0 && (D.j1 - D.j0 >= 64 || total > kHeavyItems)) {
                // (listed again for grid_heavy_kernel by the counting pass; uniform over the wave)
                if (!kFill && lane == 0) heavy_rows[atomicAdd(n_heavy, 1)] = static_cast<std::int32_t>(d);
                pos = -1;
                break;
            }
            wave_lds_fence();
            s_nq[wv][lane] = nq;
            for (int q = 0; q < kMaxLonRanges; ++q)
                if (q < nq) {
                    s_lo[wv][q][lane] = r0[q];
                    s_hi[wv][q][lane] = r1[q];
                }
            s_off[wv][lane + 1] = inc;
            if (lane == 0) s_off[wv][0] = 0;
            wave_lds_fence();
            // one grid cell per lane over the rows' cells taken together
            for (int b = 0; b < total; b += 64) {
                const int k = b + lane;
                bool hit = false;
                std::int32_t c = 0;
                if (k < total) {
                    int lo = 0, hi = 64;  // the row lane L with s_off[L] <= k < s_off[L + 1]
                    while (hi - lo > 1) {
                        const int mid = (lo + hi) >> 1;
                        if (s_off[wv][mid] <= k)
                            lo = mid;
                        else
                            hi = mid;
                    }
                    int u = k - s_off[wv][lo];
                    int a = 0;
                    const int n = s_nq[wv][lo];
                    for (int q = 0; q < n; ++q) {
                        const int qlo = s_lo[wv][q][lo], len = s_hi[wv][q][lo] - qlo + 1;
                        if (u < len) {
                            a = qlo + u;
                            break;
                        }
                        u -= len;
                    }
                    c = cell_of(G, static_cast<std::int64_t>(jb + lo) * (G.nlon - 1) + a);
                    hit = boxes_meet(G.src_box + 6 * static_cast<std::int64_t>(c), D.bx);
                }
                const std::uint64_t m = __ballot(hit);
                if (kFill && hit) {
                    const index_t w = pos + __popcll(m & ((std::uint64_t(1) << lane) - 1));
                    cand_src[w] = c;
                    pair_row[w] = static_cast<std::int32_t>(d);
                }
                pos += __popcll(m);
            }
            wave_lds_fence();
        }
        if (!kFill && lane == 0 && pos >= 0) count[d] = pos;
    }
}

// block-wide inclusive prefix sum of one int per thread (every thread calls it)
template <int BLOCK>
__device__ inline int block_inclusive_scan(int v, int *s_part) {
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    for (int o = 1; o < 64; o <<= 1) {
        const int t = __shfl_up(v, o, 64);
        if (lane >= o) v += t;
    }
    __syncthreads();  // (s_part may still be read by a previous scan)
    if (lane == 63) s_part[wv] = v;
    __syncthreads();
    for (int w = 0; w < wv; ++w) v += s_part[w];
    return v;
}

// One workgroup of kHeavyBlock threads per listed cell of more than 64 rows
// or more than kHeavyItems grid cells (the cells over and around the source
// grid's poles: at most 10,800 grid cells at 0.1 deg -> C768).  Per group of
// kHeavyBlock rows: one row per thread (ranges and counts in LDS, their
// prefix scan); then passes of up to kHeavyChunks 64-cell chunks: the 16
// waves test the chunks in turn, keeping each chunk's hit mask in LDS; a
// prefix scan of the chunks' hit counts places each chunk, and the filling
// pass writes the hits in row, range and column order, as grid_long_kernel.
constexpr int kHeavyBlock = 1024;
constexpr int kHeavyWaves = kHeavyBlock / 64;
constexpr int kHeavyChunks = kHeavyBlock;  // chunks per pass: one per thread for the scan
constexpr int kHeavyGrid = 256;

template <bool kFill>
__global__ void __launch_bounds__(kHeavyBlock) grid_heavy_kernel(GridView G, const float *dst_box, const std::int32_t *heavy_rows, const int *n_heavy,
                                                                 index_t *count, const index_t *cand_off, std::int32_t *cand_src,
                                                                 std::int32_t *pair_row) {
    extern __shared__ double s_tab[];
    __shared__ int s_off[kHeavyBlock + 1];  // grid cells before row thread t
    __shared__ int s_lo[kMaxLonRanges][kHeavyBlock], s_hi[kMaxLonRanges][kHeavyBlock];
    __shared__ int s_nq[kHeavyBlock];
    __shared__ unsigned long long s_mask[kHeavyChunks];
    __shared__ int s_cpre[kHeavyChunks + 1];  // hits before chunk c of the pass
    __shared__ int s_part[kHeavyWaves];
    const Tables T = load_tables(G, s_tab);
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int nh = *n_heavy;
    // (global cell k of the row group -> its source cell)
    auto cell_at_k = [&](int jb, int k) {
        int lo = 0, hi = kHeavyBlock;  // the row thread L with s_off[L] <= k < s_off[L + 1]
        while (hi - lo > 1) {
            const int mid = (lo + hi) >> 1;
            if (s_off[mid] <= k)
                lo = mid;
            else
                hi = mid;
        }
        int u = k - s_off[lo], a = 0;
        const int n = s_nq[lo];
        for (int q = 0; q < n; ++q) {
            const int qlo = s_lo[q][lo], len = s_hi[q][lo] - qlo + 1;
            if (u < len) {
                a = qlo + u;
                break;
            }
            u -= len;
        }
        return cell_of(G, static_cast<std::int64_t>(jb + lo) * (G.nlon - 1) + a);
    };
    for (int hc = blockIdx.x; hc < nh; hc += gridDim.x) {  // (uniform over the workgroup)
        const index_t d = heavy_rows[hc];
        Dst D;
        dst_setup(G, T, dst_box, d, D);
        index_t pos = kFill ? cand_off[d] : 0;
        for (int jb = D.j0; jb <= D.j1; jb += kHeavyBlock) {
            int r0[kMaxLonRanges], r1[kMaxLonRanges];
            int nq = 0, cnt = 0;
            if (jb + tid <= D.j1) {
                nq = row_ranges(G, T, jb + tid, D, r0, r1);
                for (int q = 0; q < nq; ++q) cnt += r1[q] - r0[q] + 1;
            }
            s_nq[tid] = nq;
            for (int q = 0; q < kMaxLonRanges; ++q)
                if (q < nq) {
                    s_lo[q][tid] = r0[q];
                    s_hi[q][tid] = r1[q];
                }
            const int inc = block_inclusive_scan<kHeavyBlock>(cnt, s_part);
            s_off[tid + 1] = inc;
            if (tid == 0) s_off[0] = 0;
            __syncthreads();
            const int total = s_off[kHeavyBlock];
            for (int cb = 0; cb < total; cb += 64 * kHeavyChunks) {
                const int nch = min(kHeavyChunks, (total - cb + 63) / 64);
                for (int c = wv; c < nch; c += kHeavyWaves) {
                    const int k = cb + 64 * c + lane;
                    bool hit = false;
                    if (k < total) hit = boxes_meet(G.src_box + 6 * static_cast<std::int64_t>(cell_at_k(jb, k)), D.bx);
                    const std::uint64_t m = __ballot(hit);
                    if (lane == 0) s_mask[c] = m;
                }
                __syncthreads();
                const int hits = tid < nch ? __popcll(s_mask[tid]) : 0;
                const int hinc = block_inclusive_scan<kHeavyBlock>(hits, s_part);
                s_cpre[tid + 1] = hinc;
                if (tid == 0) s_cpre[0] = 0;
                __syncthreads();
                if (kFill)
                    for (int c = wv; c < nch; c += kHeavyWaves) {
                        const std::uint64_t m = s_mask[c];
                        if ((m >> lane) & 1u) {
                            const index_t w = pos + s_cpre[c] + __popcll(m & ((std::uint64_t(1) << lane) - 1));
                            cand_src[w] = cell_at_k(jb, cb + 64 * c + lane);
                            pair_row[w] = static_cast<std::int32_t>(d);
                        }
                    }
                pos += s_cpre[nch];
                __syncthreads();  // (s_mask and s_cpre are rewritten by the next pass)
            }
        }
        if (!kFill && tid == 0) count[d] = pos;
    }
}

__global__ void fill_u32_kernel(unsigned *p, std::int64_t n, unsigned v) {
    const std::int64_t i = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

unsigned count_blocks(index_t n_dst, int block) {
    return static_cast<unsigned>(std::clamp<index_t>((n_dst + block - 1) / block, 1, kCountBlocks * kQueryBlock / block));
}
std::size_t rows_lds(int nrow) { return sizeof(double) * ((kRowFields * static_cast<std::size_t>(nrow) + 1) / 2); }
bool lon_fits(int nlon, int nrow) { return rows_lds(nrow) + sizeof(double) * nlon <= kMaxLds; }
std::size_t query_lds(int nlon, int nrow) { return rows_lds(nrow) + (lon_fits(nlon, nrow) ? sizeof(double) * nlon : 0); }

}  // namespace

bool grid_index_build(const Mesh &src, const GridAxes &ax, const float *src_boxes, GridIndex &G, hipStream_t s) {
    PhaseTrace tr("broad_phase", s);
    G.ready = false;
    const int nlon = static_cast<int>(ax.ulon.size()), nlat = static_cast<int>(ax.ulat.size());
    const index_t n_cells = src.n_cells(), n_nodes = src.n_nodes();
    if (nlon < 2 || nlat < 2 || static_cast<index_t>(nlon - 1) * (nlat - 1) != n_cells || n_cells >= (index_t(1) << 31)) return false;
    if (src.coord_system() != CoordinateSystem::SphericalDeg || src.min_cell_vertices() != 4 || src.max_cell_vertices() != 4) return false;
    // the axes (node classes) and the query tables live in LDS
    if (sizeof(double) * (static_cast<std::size_t>(nlon) + nlat) > kMaxLds || rows_lds(nlat - 1) > kMaxLds) return false;
    const int rows = nlat - 1;
    G.nlon = nlon;
    G.nlat = nlat;
    G.axes.resize(static_cast<std::size_t>(nlon + nlat));
    PinnedArray<double> h_axes;
    h_axes.resize(static_cast<std::size_t>(nlon + nlat));
    std::copy(ax.ulon.begin(), ax.ulon.end(), h_axes.data());
    std::copy(ax.ulat.begin(), ax.ulat.end(), h_axes.data() + nlon);
    G.ulon0 = ax.ulon.front();
    G.dlon = (ax.ulon.back() - ax.ulon.front()) / (nlon - 1);
    if (!(G.dlon > 0.0)) return false;
    G.axes.upload(h_axes.data(), h_axes.size(), s);
    G.cell_at.resize(static_cast<std::size_t>(n_cells));
    G.rows.resize(kRowFields * static_cast<std::size_t>(rows));
    DeviceBuffer<std::int32_t> na(std::max<index_t>(n_nodes, 1)), nb(std::max<index_t>(n_nodes, 1));
    DeviceBuffer<unsigned> keys(kRowFields * static_cast<std::size_t>(rows));
    DeviceBuffer<int> bad(1);
    AXIS_HIP_CHECK(hipMemsetAsync(bad.get(), 0, sizeof(int), s));
    fill_u32_kernel<<<grid_for(n_cells, 256), 256, 0, s>>>(reinterpret_cast<unsigned *>(G.cell_at.get()), n_cells, 0xFFFFFFFFu);
    AXIS_LAUNCH_CHECK();
    // minima start at the largest key, maxima at zero: zlo min, zhi max, overhang max, rlo min, rhi max
    fill_u32_kernel<<<grid_for(rows, 256), 256, 0, s>>>(keys.get(), rows, 0xFFFFFFFFu);
    AXIS_LAUNCH_CHECK();
    AXIS_HIP_CHECK(hipMemsetAsync(keys.get() + rows, 0, 2 * sizeof(unsigned) * rows, s));
    fill_u32_kernel<<<grid_for(rows, 256), 256, 0, s>>>(keys.get() + 3 * rows, rows, 0xFFFFFFFFu);
    AXIS_LAUNCH_CHECK();
    AXIS_HIP_CHECK(hipMemsetAsync(keys.get() + 4 * rows, 0, sizeof(unsigned) * rows, s));
    node_class_kernel<<<grid_for(n_nodes, 1024), 1024, sizeof(double) * (nlon + nlat), s>>>(n_nodes, src.d_coords.get(), src.d_coords.get() + n_nodes,
                                                                                         G.axes.get(), nlon, nlat, na.get(), nb.get());
    AXIS_LAUNCH_CHECK();
    cell_grid_kernel<<<grid_for(n_cells, 256), 256, 0, s>>>(n_cells, src.d_offsets.get(), src.d_indices.get(), na.get(), nb.get(), G.axes.get(), nlon,
                                                            nlat, src_boxes, G.cell_at.get(), bad.get(), keys.get());
    AXIS_LAUNCH_CHECK();
    row_finish_kernel<<<1, 64, 0, s>>>(rows, keys.get(), G.rows.get());
    AXIS_LAUNCH_CHECK();
    PinnedArray<int> h_bad;
    h_bad.resize(1);
    AXIS_HIP_CHECK(hipMemcpyAsync(h_bad.data(), bad.get(), sizeof(int), hipMemcpyDeviceToHost, s));
    AXIS_HIP_CHECK(hipStreamSynchronize(s));
    G.src_boxes = src_boxes;
    G.ready = (h_bad[0] & 1) == 0;
    G.grid_order = (h_bad[0] & 2) == 0;
    tr.mark("grid_map");
    return G.ready;
}

index_t grid_query(const GridIndex &G, const float *dst_boxes, index_t n_dst, DeviceBuffer<index_t> &cand_off, DeviceBuffer<std::int32_t> &cand_src,
                   DeviceBuffer<std::int32_t> &pair_row, hipStream_t s) {
    if (!G.ready) throw InvalidArgument("grid_query: no grid index");
    PhaseTrace tr("broad_phase", s);
    const double u0 = G.ulon0, du = 1.0 / G.dlon;
    const GridView V{G.axes.get(), G.rows.get(), G.nlon, G.nlat - 1, lon_fits(G.nlon, G.nlat - 1), u0, du, G.grid_order ? nullptr : G.cell_at.get(),
                     G.src_boxes};
    const std::size_t lds = query_lds(G.nlon, G.nlat - 1);
    const index_t nd1 = std::max<index_t>(n_dst, 1);
    cand_off.resize(n_dst + 1);
    DeviceBuffer<index_t> count(nd1);
    DeviceBuffer<std::int32_t> long_rows(nd1);
    DeviceBuffer<std::int32_t> heavy_rows(nd1);
    DeviceBuffer<int> n_lists(2);  // listed cells, of which heavy
    int *n_long = n_lists.get(), *n_heavy = n_lists.get() + 1;
    AXIS_HIP_CHECK(hipMemsetAsync(n_lists.get(), 0, 2 * sizeof(int), s));
    if (n_dst > 0) {
        // counting: 1024-thread workgroups when the tables leave room for
        // fewer than five 256-thread workgroups per CU (more waves resident)
        if (lds > kBigTables)
            grid_count_kernel<kLongBlock><<<count_blocks(n_dst, kLongBlock), kLongBlock, lds, s>>>(V, dst_boxes, n_dst, count.get(),
                                                                                                  long_rows.get(), n_long);
        else
            grid_count_kernel<kQueryBlock><<<count_blocks(n_dst, kQueryBlock), kQueryBlock, lds, s>>>(V, dst_boxes, n_dst, count.get(),
                                                                                                     long_rows.get(), n_long);
        AXIS_LAUNCH_CHECK();
        grid_long_kernel<false><<<kLongBlocks, kLongBlock, lds, s>>>(V, dst_boxes, long_rows.get(), n_long, count.get(), nullptr, nullptr, nullptr,
                                                                      heavy_rows.get(), n_heavy);
        AXIS_LAUNCH_CHECK();
        grid_heavy_kernel<false><<<kHeavyGrid, kHeavyBlock, lds, s>>>(V, dst_boxes, heavy_rows.get(), n_heavy, count.get(), nullptr, nullptr,
                                                                      nullptr);
        AXIS_LAUNCH_CHECK();
    }
    exclusive_scan_i64(count.get(), cand_off.get(), n_dst, s);
    PinnedArray<index_t> h_total;
    h_total.resize(1);
    AXIS_HIP_CHECK(hipMemcpyAsync(h_total.data(), cand_off.get() + n_dst, sizeof(index_t), hipMemcpyDeviceToHost, s));
    AXIS_HIP_CHECK(hipStreamSynchronize(s));
    const index_t total = h_total[0];
    tr.mark("grid_count");
    // the reference stores candidate offsets as int (weight_generator.cpp:247-254)
    if (total >= (index_t(1) << 31)) throw RuntimeError("broad phase: candidate pair count exceeds 2^31");
    cand_src.resize(std::max<index_t>(total, 1));
    pair_row.resize(std::max<index_t>(total, 1));
    if (n_dst > 0 && total > 0) {
        // the window beside the tables: 256-thread workgroups with small
        // tables, else 1024-thread ones (as the count)
        const unsigned tab = static_cast<unsigned>(lds);
        if (lds > kBigTables)
            grid_fill_kernel<kLongBlock, kFillWinBig><<<count_blocks(n_dst, kLongBlock), kLongBlock, lds + fill_lds(kLongBlock, kFillWinBig), s>>>(
                V, dst_boxes, n_dst, cand_off.get(), cand_src.get(), pair_row.get(), tab);
        else
            grid_fill_kernel<kQueryBlock, kFillWin><<<count_blocks(n_dst, kQueryBlock), kQueryBlock, lds + fill_lds(kQueryBlock, kFillWin), s>>>(
                V, dst_boxes, n_dst, cand_off.get(), cand_src.get(), pair_row.get(), tab);
        AXIS_LAUNCH_CHECK();
        grid_long_kernel<true><<<kLongBlocks, kLongBlock, lds, s>>>(V, dst_boxes, long_rows.get(), n_long, nullptr, cand_off.get(), cand_src.get(),
                                                                     pair_row.get(), nullptr, nullptr);
        AXIS_LAUNCH_CHECK();
        grid_heavy_kernel<true><<<kHeavyGrid, kHeavyBlock, lds, s>>>(V, dst_boxes, heavy_rows.get(), n_heavy, nullptr, cand_off.get(),
                                                                     cand_src.get(), pair_row.get());
        AXIS_LAUNCH_CHECK();
    }
    tr.mark("grid_fill");
    return total;
}

}  // namespace axis_hip::detail
