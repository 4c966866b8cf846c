// Analytic weight paths for structured lat-lon pairs:
//
//   * conservative_regular      generate_conservative_rect
//                               (weight_generator_conservative_rect.cpp:87-395):
//                               both meshes regular; overlaps are exact
//                               lat-lon rectangle intersections
//                               dlon * (sin(hi) - sin(lo));
//   * conservative_rectilinear  generate_conservative_rect_nonuniform
//                               (weight_generator_conservative_rect_nonuniform.cpp:52-313):
//                               both meshes rectilinear, the same overlaps
//                               over the unique node longitudes / latitudes;
//   * bilinear_regular          generate_bilinear_rect
//                               (weight_generator_bilinear_rect.cpp:33-266):
//                               2x2 cell-centred stencils by index arithmetic.
//
// Every per-column and per-row decision (the normalised longitude edges,
// source index ranges, latitude edges) depends on one destination column or
// row only, so the host computes those few arrays with the reference's
// arithmetic, and the sines of every latitude edge with host libm (the
// reference evaluates std::sin on the host).  The kernels then do exact
// double + - * / on them: weights, fractions and areas equal the reference's
// bit for bit.  One thread per destination cell counts its entries and sums
// frac_b in the reference's (js, is) order, a second pass writes them in
// column order; frac_a sums each source column's ov / area_src in row order
// after a stable column sort.
#include <algorithm>
#include <chrono>
#include <cmath>

#include "axis_hip.hpp"
#include "host_util.hpp"

namespace axis_hip::detail {
namespace {

constexpr double kD2R = 3.14159265358979323846 / 180.0;  // the reference's folded deg2rad constant

struct RectArgs {
    int path;  // 0 regular, 1 rectilinear
    int cs;    // 0 SphericalDeg, 1 SphericalRad, 2 Cartesian3D
    int s_glob;
    index_t sni, snj, dni, dnj;
    double s_lon0, s_dlon, s_lat0, s_dlat;  // regular source edges
    const double *s_ulon, *s_ulat;          // rectilinear source edges
    // sines of the lower / upper latitude edge of every source and
    // destination row (radian argument)
    const double *s_sin_lo, *s_sin_hi, *d_sin_lo, *d_sin_hi;
    // per destination column: normalised edges, source index range, the
    // first index of the column-order walk, in-loop area longitude width
    const double *c_nlx, *c_nhx;
    const int *c_is0, *c_is1, *c_k0;
    // per destination row: latitude edges and source row range
    const double *r_dly, *r_dhy;
    const int *r_js0, *r_js1;
    const double *s_area;  // regular: final source areas; rectilinear: mesh areas or null
    const double *d_area;  // regular: final destination areas; rectilinear: mesh areas or null
    const std::int32_t *s_mask, *d_mask;
};

__device__ inline int wrap_index(const RectArgs &a, int is) {
    if (!a.s_glob) return is;
    const int n = static_cast<int>(a.sni);
    return ((is % n) + n) % n;
}

__device__ inline void src_edges(const RectArgs &a, int is, int ia, int js, double &slx, double &shx, double &sly, double &shy) {
    if (a.path == 0) {
        slx = a.s_lon0 + static_cast<double>(is) * a.s_dlon;
        shx = slx + a.s_dlon;
        sly = a.s_lat0 + static_cast<double>(js) * a.s_dlat;
        shy = sly + a.s_dlat;
    } else {
        slx = a.s_ulon[ia];
        shx = a.s_ulon[ia + 1];
        if (a.s_glob && is >= static_cast<int>(a.sni)) {
            slx += 360.0;
            shx += 360.0;
        }
        sly = a.s_ulat[js];
        shy = a.s_ulat[js + 1];
    }
}

// rect_overlap_spherical / rect_overlap (…_rect.cpp:40-69): the sines come
// from the tables; fmax / fmin pick an edge, and equal edges share a sine.
__device__ inline double overlap(const RectArgs &a, double slx, double shx, double sly, double shy, int js, int jd, double nlx, double nhx,
                                 double dly, double dhy) {
    double dx = fmax(0.0, fmin(shx, nhx) - fmax(slx, nlx));
    if (a.cs == 2) return dx * fmax(0.0, fmin(shy, dhy) - fmax(sly, dly));
    const double lo = fmax(sly, dly), hi = fmin(shy, dhy);
    if (hi <= lo || dx <= 0.0) return 0.0;
    const double s_lo = sly >= dly ? a.s_sin_lo[js] : a.d_sin_lo[jd];
    const double s_hi = shy <= dhy ? a.s_sin_hi[js] : a.d_sin_hi[jd];
    if (a.cs == 0) dx *= kD2R;
    return dx * (s_hi - s_lo);
}

__device__ inline double band(const RectArgs &a, double dlon, double sin_lo, double sin_hi, double lo, double hi) {
    if (a.cs == 2) return dlon * (hi - lo);
    if (a.cs == 0) dlon *= kD2R;
    return dlon * (sin_hi - sin_lo);
}

// The destination area the weights divide by: the regular path's final
// array; the rectilinear path's mesh area or the band of the normalised
// longitude edges (…_nonuniform.cpp:116-135).
__device__ inline double dst_area(const RectArgs &a, index_t j, int id, int jd) {
    if (a.path == 0 || a.d_area) return a.d_area[j];
    return band(a, a.c_nhx[id] - a.c_nlx[id], a.d_sin_lo[jd], a.d_sin_hi[jd], a.r_dly[jd], a.r_dhy[jd]);
}

// One entry's overlap and source area, or false when the reference skips it.
__device__ inline bool entry(const RectArgs &a, index_t j, int is, int js, int id, int jd, index_t &i, double &ov, double &as) {
    const int ia = wrap_index(a, is);
    i = static_cast<index_t>(js) * a.sni + ia;
    if (a.s_mask && a.s_mask[i] == 0) return false;
    double slx, shx, sly, shy;
    src_edges(a, is, ia, js, slx, shx, sly, shy);
    if (a.path == 0) {
        as = a.s_area[i];
        if (as <= 0.0) return false;
        ov = overlap(a, slx, shx, sly, shy, js, jd, a.c_nlx[id], a.c_nhx[id], a.r_dly[jd], a.r_dhy[jd]);
        return ov > 0.0;
    }
    ov = overlap(a, slx, shx, sly, shy, js, jd, a.c_nlx[id], a.c_nhx[id], a.r_dly[jd], a.r_dhy[jd]);
    if (!(ov > 1e-12)) return false;
    as = a.s_area ? a.s_area[i] : band(a, shx - slx, a.s_sin_lo[js], a.s_sin_hi[js], sly, shy);
    return as > 0.0;
}

// Pass 1: entry count and frac_b (unclamped) in the reference's order.
__global__ void rect_count_kernel(RectArgs a, index_t *__restrict__ cnt, double *__restrict__ fb) {
    const index_t j = static_cast<index_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (j >= a.dni * a.dnj) return;
    const int id = static_cast<int>(j % a.dni), jd = static_cast<int>(j / a.dni);
    index_t c = 0;
    double f = 0.0;
    const double ad = dst_area(a, j, id, jd);
    if (!(a.d_mask && a.d_mask[j] == 0) && ad > 0.0) {
        for (int js = a.r_js0[jd]; js <= a.r_js1[jd]; ++js)
            for (int is = a.c_is0[id]; is <= a.c_is1[id]; ++is) {
                index_t i;
                double ov, as;
                if (!entry(a, j, is, js, id, jd, i, ov, as)) continue;
                ++c;
                f += ov / ad;
            }
    }
    cnt[j] = c;
    fb[j] = f;
}

// Pass 2: the entries in column order (rows ascending, then the column walk
// that starts at the wrapped index 0), weights normalised as configured.
__global__ void rect_fill_kernel(RectArgs a, int norm, const index_t *__restrict__ row_ptr, const double *__restrict__ fb, index_t *__restrict__ col,
                                 double *__restrict__ val, double *__restrict__ part) {
    const index_t j = static_cast<index_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (j >= a.dni * a.dnj) return;
    index_t o = row_ptr[j];
    if (o == row_ptr[j + 1]) return;
    const int id = static_cast<int>(j % a.dni), jd = static_cast<int>(j / a.dni);
    const double ad = dst_area(a, j, id, jd);
    const double f = fb[j];
    const int is0 = a.c_is0[id], is1 = a.c_is1[id], k0 = a.c_k0[id];
    const int n = is1 - is0 + 1;
    for (int js = a.r_js0[jd]; js <= a.r_js1[jd]; ++js)
        for (int t = 0; t < n; ++t) {
            const int is = k0 + t <= is1 ? k0 + t : is0 + (k0 + t - is1 - 1);
            index_t i;
            double ov, as;
            if (!entry(a, j, is, js, id, jd, i, ov, as)) continue;
            double w = ov / ad;
            if (a.path == 0) w = fmax(w, 0.0);
            if (norm == 1 && f > 0.0) w /= f;
            col[o] = i;
            val[o] = w;
            part[o] = ov / as;
            ++o;
        }
}

__global__ void clamp_kernel(double *__restrict__ p, index_t n) {
    const index_t i = static_cast<index_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) p[i] = fmin(p[i], 1.0);
}

// Final areas (…_rect.cpp:208-252; …_nonuniform.cpp:253-297): the mesh's
// areas, else the band of the cell's column width and row edges.
__global__ void rect_area_kernel(int cs, index_t ni, index_t nj, const double *__restrict__ width, const double *__restrict__ row_lo,
                                 const double *__restrict__ row_hi, const double *__restrict__ sin_lo, const double *__restrict__ sin_hi,
                                 const double *__restrict__ given, double *__restrict__ out) {
    const index_t c = static_cast<index_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (c >= ni * nj) return;
    if (given) {
        out[c] = given[c];
        return;
    }
    const index_t jj = c / ni, ii = c % ni;
    double dlon = width[ii];
    if (cs == 2) {
        out[c] = dlon * (row_hi[jj] - row_lo[jj]);
        return;
    }
    if (cs == 0) dlon *= kD2R;
    out[c] = dlon * (sin_hi[jj] - sin_lo[jj]);
}

// frac_a: Σ part over each source column's entries in row order (the CSR
// is row-ordered and the column sort stable), clamped to 1.
__global__ void column_keys_kernel(const index_t *__restrict__ col, index_t nnz, std::uint32_t *__restrict__ key, std::uint32_t *__restrict__ perm) {
    const index_t e = static_cast<index_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (e >= nnz) return;
    key[e] = static_cast<std::uint32_t>(col[e]);
    perm[e] = static_cast<std::uint32_t>(e);
}

__global__ void column_sum_kernel(const std::uint32_t *__restrict__ key, const std::uint32_t *__restrict__ perm, index_t nnz,
                                  const double *__restrict__ part, double *__restrict__ out) {
    const index_t e = static_cast<index_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (e >= nnz) return;
    const std::uint32_t c = key[e];
    if (e > 0 && key[e - 1] == c) return;
    double s = 0.0;
    for (index_t k = e; k < nnz && key[k] == c; ++k) s += part[perm[k]];
    out[c] = fmin(s, 1.0);
}

// row-length histogram (lengths 0 .. 16, then longer) in stats[0 .. 17] and
// the longest row in stats[18]
__global__ void __launch_bounds__(256) row_stats_kernel(const index_t *__restrict__ row_ptr, index_t n_rows, unsigned long long *__restrict__ stats) {
    constexpr int kMax = 16;
    __shared__ unsigned hist[kMax + 2];
    __shared__ unsigned long long longest;
    if (threadIdx.x < kMax + 2) hist[threadIdx.x] = 0;
    if (threadIdx.x == 0) longest = 0;
    __syncthreads();
    const index_t r = static_cast<index_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (r < n_rows) {
        const index_t len = row_ptr[r + 1] - row_ptr[r];
        atomicAdd(&hist[len > kMax ? kMax + 1 : len], 1u);
        atomicMax(&longest, static_cast<unsigned long long>(len));
    }
    __syncthreads();
    if (threadIdx.x < kMax + 2 && hist[threadIdx.x]) atomicAdd(&stats[threadIdx.x], static_cast<unsigned long long>(hist[threadIdx.x]));
    if (threadIdx.x == 0 && longest) atomicMax(&stats[kMax + 2], longest);
}

__global__ void zero_kernel(double *__restrict__ p, index_t n) {
    const index_t i = static_cast<index_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0.0;
}

// ------------------------------------------------------------ bilinear
// generate_bilinear_rect (weight_generator_bilinear_rect.cpp:33-266).
struct BilArgs {
    index_t ni, nj;
    double lon0, lat0, dlon, dlat;
    int periodic;
};

// The reference's compute_weights for one destination cell: false when the
// cell's planar centroid is outside the grid.
__device__ inline bool bil_weights(const BilArgs &g, double lon, double lat, index_t *ix, double *w) {
    if (g.periodic) {
        double off = lon - g.lon0;
        off = off - floor(off / 360.0) * 360.0;
        lon = g.lon0 + off;
    }
    const double fi = (lon - g.lon0) / g.dlon - 0.5, fj = (lat - g.lat0) / g.dlat - 0.5;
    const double tol = 1.0e-9;
    if (fj < -0.5 - tol || fj > static_cast<double>(g.nj) - 0.5 + tol) return false;
    if (!g.periodic && (fi < -0.5 - tol || fi > static_cast<double>(g.ni) - 0.5 + tol)) return false;
    const int ni = static_cast<int>(g.ni), nj = static_cast<int>(g.nj);
    int i = static_cast<int>(floor(fi)), j = static_cast<int>(floor(fj));
    if (j < 0) j = 0;
    if (j >= nj - 1) j = nj - 2;
    if (j < 0) j = 0;
    if (!g.periodic) {
        if (i < 0) i = 0;
        if (i >= ni - 1) i = ni - 2;
        if (i < 0) i = 0;
    } else {
        i = ((i % ni) + ni) % ni;
    }
    double tx = fi - static_cast<double>(i), ty = fj - static_cast<double>(j);
    tx = fmin(fmax(tx, 0.0), 1.0);
    ty = fmin(fmax(ty, 0.0), 1.0);
    int i1 = i + 1, j1 = j + 1;
    if (i1 >= ni) i1 = g.periodic ? 0 : ni - 1;
    if (j1 >= nj) j1 = nj - 1;
    ix[0] = static_cast<index_t>(j) * g.ni + i;
    ix[1] = static_cast<index_t>(j) * g.ni + i1;
    ix[2] = static_cast<index_t>(j1) * g.ni + i;
    ix[3] = static_cast<index_t>(j1) * g.ni + i1;
    w[0] = (1.0 - tx) * (1.0 - ty);
    w[1] = tx * (1.0 - ty);
    w[2] = (1.0 - tx) * ty;
    w[3] = tx * ty;
    return true;
}

// Planar mean of the cell's nodes (…_bilinear_rect.cpp:205-217).
__device__ inline void planar_centroid(const index_t *off, const index_t *idx, const double *coords, index_t n_nodes, index_t c, double &lon,
                                       double &lat) {
    const index_t b = off[c], e = off[c + 1];
    double sx = 0.0, sy = 0.0;
    for (index_t k = b; k < e; ++k) {
        sx += coords[idx[k]];
        sy += coords[n_nodes + idx[k]];
    }
    lon = sx / static_cast<double>(e - b);
    lat = sy / static_cast<double>(e - b);
}

__global__ void bil_count_kernel(BilArgs g, const index_t *__restrict__ off, const index_t *__restrict__ idx, const double *__restrict__ coords,
                                 index_t n_nodes, index_t n_dst, index_t *__restrict__ cnt, unsigned long long *__restrict__ first_unmapped) {
    const index_t c = static_cast<index_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (c >= n_dst) return;
    double lon, lat;
    planar_centroid(off, idx, coords, n_nodes, c, lon, lat);
    index_t ix[4];
    double w[4];
    const bool ok = bil_weights(g, lon, lat, ix, w);
    cnt[c] = ok ? 4 : 0;
    if (!ok) atomicMin(first_unmapped, static_cast<unsigned long long>(c));
}

// The four entries sorted by column (stable: equal columns keep the
// reference's corner order).
__global__ void bil_fill_kernel(BilArgs g, const index_t *__restrict__ off, const index_t *__restrict__ idx, const double *__restrict__ coords,
                                index_t n_nodes, index_t n_dst, const index_t *__restrict__ row_ptr, index_t *__restrict__ col, double *__restrict__ val,
                                double *__restrict__ frac_b) {
    const index_t c = static_cast<index_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (c >= n_dst) return;
    const index_t o = row_ptr[c];
    if (row_ptr[c + 1] == o) {
        frac_b[c] = 0.0;
        return;
    }
    double lon, lat;
    planar_centroid(off, idx, coords, n_nodes, c, lon, lat);
    index_t ix[4];
    double w[4];
    bil_weights(g, lon, lat, ix, w);
    for (int k = 1; k < 4; ++k)
        for (int m = k; m > 0 && ix[m - 1] > ix[m]; --m) {
            const index_t ti = ix[m];
            ix[m] = ix[m - 1];
            ix[m - 1] = ti;
            const double tw = w[m];
            w[m] = w[m - 1];
            w[m - 1] = tw;
        }
    for (int k = 0; k < 4; ++k) {
        col[o + k] = ix[k];
        val[o + k] = w[k];
    }
    frac_b[c] = 1.0;
}

__global__ void fill_value_kernel(double *__restrict__ p, index_t n, double v) {
    const index_t i = static_cast<index_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// ------------------------------------------------------------------ host
int cs_code(const Mesh &m) {
    return m.coord_system() == CoordinateSystem::SphericalDeg ? 0 : (m.coord_system() == CoordinateSystem::SphericalRad ? 1 : 2);
}

double host_sin(double x, int cs) { return std::sin(cs == 0 ? x * kD2R : x); }

template <class T>
const T *upload(DeviceBuffer<T> &d, const std::vector<T> &h, hipStream_t s) {
    d.resize(std::max<std::size_t>(h.size(), 1));
    d.upload(h.data(), h.size(), s);
    return d.get();
}

// Host-side tables of one structured pair.
struct RectTables {
    std::vector<double> c_nlx, c_nhx, r_dly, r_dhy, s_sin_lo, s_sin_hi, d_sin_lo, d_sin_hi;
    std::vector<int> c_is0, c_is1, c_k0, r_js0, r_js1;
    // final areas: per-column widths and row edges of each side
    std::vector<double> s_width, d_width, s_row_lo, s_row_hi, d_row_lo, d_row_hi;
};

// The column-order walk over [is0, is1]: starts at the first index that
// wraps to column 0 (global sources), else at is0.  A range longer than the
// source's columns would repeat columns and is refused.
int walk_start(int is0, int is1, index_t sni, bool glob) {
    if (!glob || is1 < is0) return is0;
    if (static_cast<index_t>(is1) - is0 + 1 > sni) throw InvalidArgument("rectangle fast path: a destination cell spans more than the source's longitudes");
    const int n = static_cast<int>(sni);
    for (int is = is0 + 1; is <= is1; ++is)
        if (((is % n) + n) % n == 0) return is;
    return is0;
}

std::unique_ptr<InterpolationMatrix> assemble(const Mesh &src, const Mesh &dst, const RegridConfig &cfg, RectArgs a, const RectTables &t,
                                              hipStream_t s, std::chrono::steady_clock::time_point t0) {
    const index_t n_src = src.n_cells(), n_dst = dst.n_cells();
    // every table in one double and one int upload
    const std::vector<double> *dtab[] = {&t.c_nlx,    &t.c_nhx,    &t.r_dly,   &t.r_dhy,    &t.s_sin_lo, &t.s_sin_hi, &t.d_sin_lo,
                                         &t.d_sin_hi, &t.s_width, &t.s_row_lo, &t.s_row_hi, &t.d_width, &t.d_row_lo, &t.d_row_hi};
    const std::vector<int> *itab[] = {&t.c_is0, &t.c_is1, &t.c_k0, &t.r_js0, &t.r_js1};
    std::vector<double> hd;
    std::vector<int> hi;
    std::size_t od[14], oi[5];
    for (int k = 0; k < 14; ++k) {
        od[k] = hd.size();
        hd.insert(hd.end(), dtab[k]->begin(), dtab[k]->end());
    }
    for (int k = 0; k < 5; ++k) {
        oi[k] = hi.size();
        hi.insert(hi.end(), itab[k]->begin(), itab[k]->end());
    }
    DeviceBuffer<double> bd(std::max<std::size_t>(hd.size(), 1));
    DeviceBuffer<int> bi(std::max<std::size_t>(hi.size(), 1));
    if (!hd.empty()) bd.upload(hd.data(), hd.size(), s);
    if (!hi.empty()) bi.upload(hi.data(), hi.size(), s);
    auto D = [&](int k) -> const double * { return bd.get() + od[k]; };
    auto I = [&](int k) -> const int * { return bi.get() + oi[k]; };
    a.c_nlx = D(0);
    a.c_nhx = D(1);
    a.r_dly = D(2);
    a.r_dhy = D(3);
    a.s_sin_lo = D(4);
    a.s_sin_hi = D(5);
    a.d_sin_lo = D(6);
    a.d_sin_hi = D(7);
    a.c_is0 = I(0);
    a.c_is1 = I(1);
    a.c_k0 = I(2);
    a.r_js0 = I(3);
    a.r_js1 = I(4);
    a.s_mask = src.has_mask() ? src.d_mask.get() : nullptr;
    a.d_mask = dst.has_mask() ? dst.d_mask.get() : nullptr;
    auto M = std::make_unique<InterpolationMatrix>();
    M->n_src = n_src;
    M->n_dst = n_dst;
    const index_t ns = std::max<index_t>(n_src, 1), nd = std::max<index_t>(n_dst, 1);
    M->frac_a.resize(ns);
    M->frac_b.resize(nd);
    M->area_a.resize(ns);
    M->area_b.resize(nd);
    M->row_ptr.resize(n_dst + 1);
    // final areas first: the regular path's weights divide by them
    rect_area_kernel<<<grid_for(ns, 256), 256, 0, s>>>(a.cs, a.sni, a.snj, D(8), D(9), D(10), a.s_sin_lo, a.s_sin_hi,
                                                       src.has_areas() ? src.d_areas.get() : nullptr, M->area_a.get());
    rect_area_kernel<<<grid_for(nd, 256), 256, 0, s>>>(a.cs, a.dni, a.dnj, D(11), D(12), D(13), a.d_sin_lo, a.d_sin_hi,
                                                       dst.has_areas() ? dst.d_areas.get() : nullptr, M->area_b.get());
    AXIS_LAUNCH_CHECK();
    if (a.path == 0) {
        a.s_area = M->area_a.get();
        a.d_area = M->area_b.get();
    } else {
        a.s_area = src.has_areas() ? src.d_areas.get() : nullptr;
        a.d_area = dst.has_areas() ? dst.d_areas.get() : nullptr;
    }
    DeviceBuffer<index_t> cnt(nd + 1);
    DeviceBuffer<double> fb(nd);
    rect_count_kernel<<<grid_for(nd, 256), 256, 0, s>>>(a, cnt.get(), fb.get());
    AXIS_LAUNCH_CHECK();
    exclusive_scan_i64(cnt.get(), M->row_ptr.get(), n_dst, s);
    index_t nnz = 0;
    AXIS_HIP_CHECK(hipMemcpyAsync(&nnz, M->row_ptr.get() + n_dst, sizeof(index_t), hipMemcpyDeviceToHost, s));
    AXIS_HIP_CHECK(hipStreamSynchronize(s));
    M->nnz = nnz;
    const index_t nz = std::max<index_t>(nnz, 1);
    M->col.resize(nz);
    M->val.resize(nz);
    DeviceBuffer<double> part(nz);
    rect_fill_kernel<<<grid_for(nd, 256), 256, 0, s>>>(a, cfg.norm_type == NormType::FracArea ? 1 : 0, M->row_ptr.get(), fb.get(), M->col.get(),
                                                       M->val.get(), part.get());
    AXIS_LAUNCH_CHECK();
    // fractions: frac_b clamped; frac_a by a stable column sort
    clamp_kernel<<<grid_for(nd, 256), 256, 0, s>>>(fb.get(), n_dst);
    AXIS_LAUNCH_CHECK();
    AXIS_HIP_CHECK(hipMemcpyAsync(M->frac_b.get(), fb.get(), sizeof(double) * n_dst, hipMemcpyDeviceToDevice, s));
    zero_kernel<<<grid_for(ns, 256), 256, 0, s>>>(M->frac_a.get(), ns);
    AXIS_LAUNCH_CHECK();
    if (nnz > 0) {
        DeviceBuffer<std::uint32_t> key(nnz), perm(nnz);
        column_keys_kernel<<<grid_for(nnz, 256), 256, 0, s>>>(M->col.get(), nnz, key.get(), perm.get());
        AXIS_LAUNCH_CHECK();
        int bits = 8;
        while (bits < 32 && (index_t(1) << bits) < n_src) bits += 8;
        radix_sort_pairs_u32(key.get(), perm.get(), static_cast<std::size_t>(nnz), bits, s);
        column_sum_kernel<<<grid_for(nnz, 256), 256, 0, s>>>(key.get(), perm.get(), nnz, part.get(), M->frac_a.get());
        AXIS_LAUNCH_CHECK();
        // the longest row and the row-length histogram (for the apply), on the device
        constexpr int kH = sizeof(M->len_hist) / sizeof(M->len_hist[0]);
        DeviceBuffer<unsigned long long> stats(kH + 1);
        AXIS_HIP_CHECK(hipMemsetAsync(stats.get(), 0, sizeof(unsigned long long) * (kH + 1), s));
        row_stats_kernel<<<grid_for(n_dst, 256), 256, 0, s>>>(M->row_ptr.get(), n_dst, stats.get());
        AXIS_LAUNCH_CHECK();
        unsigned long long h[kH + 1];
        AXIS_HIP_CHECK(hipMemcpyAsync(h, stats.get(), sizeof h, hipMemcpyDeviceToHost, s));
        AXIS_HIP_CHECK(hipStreamSynchronize(s));
        M->max_row_nnz = static_cast<int>(std::min<unsigned long long>(h[kH], 1ull << 30));
        for (int l = 0; l < kH; ++l) M->len_hist[l] = static_cast<index_t>(h[l]);
        M->len_hist_ready = true;
    }
    cell_morton_keys(dst, M->row_key, s);
    AXIS_HIP_CHECK(hipStreamSynchronize(s));  // the tables and work buffers are released on return
    M->gen_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return M;
}

}  // namespace

std::unique_ptr<InterpolationMatrix> conservative_regular(const Mesh &src, const Mesh &dst, const RegridConfig &cfg, const GridAxes &sax,
                                                          const GridAxes &dax, const MeshGeom &sg, const MeshGeom &dg, hipStream_t s) {
    const auto t0 = std::chrono::steady_clock::now();
    const int cs = cs_code(src);
    const bool deg = cs == 0;
    const double wrap = deg ? 350.0 : 6.0;
    const bool s_glob = static_cast<double>(sg.ni) * sg.delta_lon > wrap, d_glob = static_cast<double>(dg.ni) * dg.delta_lon > wrap;
    // decline: neither grid global and the destination reaches past the source (…_rect.cpp:132-143)
    if (!s_glob && !d_glob &&
        (dg.lon_min < sg.lon_min - sg.delta_lon * 0.5 || dax.ulon.back() > sax.ulon.back() + sg.delta_lon * 0.5))
        return nullptr;
    RectArgs a{};
    a.path = 0;
    a.cs = cs;
    a.s_glob = s_glob;
    a.sni = sg.ni;
    a.snj = sg.nj;
    a.dni = dg.ni;
    a.dnj = dg.nj;
    a.s_lon0 = sg.lon_min;
    a.s_dlon = sg.delta_lon;
    a.s_lat0 = sg.lat_min;
    a.s_dlat = sg.delta_lat;
    RectTables t;
    for (index_t js = 0; js < sg.nj; ++js) {
        const double lo = sg.lat_min + static_cast<double>(js) * sg.delta_lat, hi = lo + sg.delta_lat;
        t.s_row_lo.push_back(lo);
        t.s_row_hi.push_back(hi);
        t.s_sin_lo.push_back(cs == 2 ? 0.0 : host_sin(lo, cs));
        t.s_sin_hi.push_back(cs == 2 ? 0.0 : host_sin(hi, cs));
    }
    for (index_t jd = 0; jd < dg.nj; ++jd) {
        const double lo = dg.lat_min + static_cast<double>(jd) * dg.delta_lat, hi = lo + dg.delta_lat;
        t.d_row_lo.push_back(lo);
        t.d_row_hi.push_back(hi);
        t.d_sin_lo.push_back(cs == 2 ? 0.0 : host_sin(lo, cs));
        t.d_sin_hi.push_back(cs == 2 ? 0.0 : host_sin(hi, cs));
        t.r_dly.push_back(lo);
        t.r_dhy.push_back(hi);
        t.r_js0.push_back(std::max(0, static_cast<int>(std::floor((lo - sg.lat_min) / sg.delta_lat))));
        t.r_js1.push_back(std::min(static_cast<int>(sg.nj) - 1, static_cast<int>(std::floor((hi - sg.lat_min) / sg.delta_lat))));
    }
    t.s_width.assign(sg.ni, sg.delta_lon);
    t.d_width.assign(dg.ni, dg.delta_lon);
    for (index_t id = 0; id < dg.ni; ++id) {
        const double dlx = dg.lon_min + static_cast<double>(id) * dg.delta_lon, dhx = dlx + dg.delta_lon;
        double nlx = dlx, nhx = dhx;
        if (s_glob) {
            const double shift = 360.0 * std::floor((dlx - sg.lon_min) / 360.0);
            nlx = dlx - shift;
            nhx = dhx - shift;
        }
        int is0 = static_cast<int>(std::floor((nlx - sg.lon_min) / sg.delta_lon));
        int is1 = static_cast<int>(std::floor((nhx - sg.lon_min) / sg.delta_lon));
        if (!s_glob) {
            is0 = std::max(0, is0);
            is1 = std::min(static_cast<int>(sg.ni) - 1, is1);
        }
        t.c_nlx.push_back(nlx);
        t.c_nhx.push_back(nhx);
        t.c_is0.push_back(is0);
        t.c_is1.push_back(is1);
        t.c_k0.push_back(walk_start(is0, is1, sg.ni, s_glob));
    }
    return assemble(src, dst, cfg, a, t, s, t0);
}

std::unique_ptr<InterpolationMatrix> conservative_rectilinear(const Mesh &src, const Mesh &dst, const RegridConfig &cfg, const GridAxes &sax,
                                                              const GridAxes &dax, hipStream_t s) {
    const auto t0 = std::chrono::steady_clock::now();
    const int cs = cs_code(src);
    const bool deg = cs == 0;
    const std::vector<double> &sl = sax.ulon, &sa = sax.ulat, &dl = dax.ulon, &da = dax.ulat;
    const index_t sni = static_cast<index_t>(sl.size()) - 1, snj = static_cast<index_t>(sa.size()) - 1;
    const index_t dni = static_cast<index_t>(dl.size()) - 1, dnj = static_cast<index_t>(da.size()) - 1;
    const bool s_glob = sl[sni] - sl[0] > (deg ? 350.0 : 6.0);
    RectArgs a{};
    a.path = 1;
    a.cs = cs;
    a.s_glob = s_glob;
    a.sni = sni;
    a.snj = snj;
    a.dni = dni;
    a.dnj = dnj;
    DeviceBuffer<double> b_ul, b_ua;
    a.s_ulon = upload(b_ul, sl, s);
    a.s_ulat = upload(b_ua, sa, s);
    RectTables t;
    for (index_t js = 0; js < snj; ++js) {
        t.s_row_lo.push_back(sa[js]);
        t.s_row_hi.push_back(sa[js + 1]);
        t.s_sin_lo.push_back(cs == 2 ? 0.0 : host_sin(sa[js], cs));
        t.s_sin_hi.push_back(cs == 2 ? 0.0 : host_sin(sa[js + 1], cs));
    }
    for (index_t jd = 0; jd < dnj; ++jd) {
        const double lo = da[jd], hi = da[jd + 1];
        t.d_row_lo.push_back(lo);
        t.d_row_hi.push_back(hi);
        t.d_sin_lo.push_back(cs == 2 ? 0.0 : host_sin(lo, cs));
        t.d_sin_hi.push_back(cs == 2 ? 0.0 : host_sin(hi, cs));
        t.r_dly.push_back(lo);
        t.r_dhy.push_back(hi);
        const index_t b0 = std::lower_bound(sa.begin(), sa.end(), lo) - sa.begin();
        const index_t b1 = std::upper_bound(sa.begin(), sa.end(), hi) - sa.begin();
        t.r_js0.push_back(std::max(0, static_cast<int>(b0 - 1)));
        t.r_js1.push_back(std::min(static_cast<int>(snj) - 1, static_cast<int>(b1)));
    }
    for (index_t i = 0; i < sni; ++i) t.s_width.push_back(sl[i + 1] - sl[i]);
    for (index_t i = 0; i < dni; ++i) t.d_width.push_back(dl[i + 1] - dl[i]);
    for (index_t id = 0; id < dni; ++id) {
        double nlx = dl[id], nhx = dl[id + 1];
        if (s_glob) {
            const double shift = 360.0 * std::floor((nlx - sl[0]) / 360.0);
            nlx -= shift;
            nhx -= shift;
        }
        const index_t a0 = std::lower_bound(sl.begin(), sl.end(), nlx) - sl.begin();
        const index_t a1 = std::upper_bound(sl.begin(), sl.end(), nhx) - sl.begin();
        const int is0 = std::max(0, static_cast<int>(a0 - 1));
        int is1 = static_cast<int>(a1);
        if (!s_glob || nhx <= sl[sni]) is1 = std::min(static_cast<int>(sni) - 1, is1);
        t.c_nlx.push_back(nlx);
        t.c_nhx.push_back(nhx);
        t.c_is0.push_back(is0);
        t.c_is1.push_back(is1);
        t.c_k0.push_back(walk_start(is0, is1, sni, s_glob));
    }
    auto M = assemble(src, dst, cfg, a, t, s, t0);
    return M;
}

std::unique_ptr<InterpolationMatrix> bilinear_regular(const Mesh &src, const Mesh &dst, const RegridConfig &cfg, const GridAxes &sax, const MeshGeom &sg,
                                                      hipStream_t s) {
    const auto t0 = std::chrono::steady_clock::now();
    if (sg.ni == 0 || sg.nj == 0) return nullptr;
    const index_t n_src = src.n_cells(), n_dst = dst.n_cells();
    const double span = sax.ulon.back() - sg.lon_min;
    BilArgs g{sg.ni, sg.nj, sg.lon_min, sg.lat_min, sg.delta_lon, sg.delta_lat, std::abs(span - 360.0) < 1.0e-10 * sg.delta_lon ? 1 : 0};
    auto M = std::make_unique<InterpolationMatrix>();
    M->n_src = n_src;
    M->n_dst = n_dst;
    const index_t ns = std::max<index_t>(n_src, 1), nd = std::max<index_t>(n_dst, 1);
    M->frac_a.resize(ns);
    M->frac_b.resize(nd);
    M->area_a.resize(ns);
    M->area_b.resize(nd);
    M->row_ptr.resize(n_dst + 1);
    DeviceBuffer<index_t> cnt(nd + 1);
    DeviceBuffer<unsigned long long> first(1);
    const unsigned long long none = ~0ull;
    first.upload(&none, 1, s);
    bil_count_kernel<<<grid_for(nd, 256), 256, 0, s>>>(g, dst.d_offsets.get(), dst.d_indices.get(), dst.d_coords.get(), dst.n_nodes(), n_dst, cnt.get(),
                                                       first.get());
    AXIS_LAUNCH_CHECK();
    exclusive_scan_i64(cnt.get(), M->row_ptr.get(), n_dst, s);
    index_t nnz = 0;
    unsigned long long f = none;
    AXIS_HIP_CHECK(hipMemcpyAsync(&nnz, M->row_ptr.get() + n_dst, sizeof(index_t), hipMemcpyDeviceToHost, s));
    first.download(&f, 1, s);
    AXIS_HIP_CHECK(hipStreamSynchronize(s));
    if (f != none && cfg.unmapped == UnmappedAction::Error)
        throw InvalidArgument("Unmapped destination cell " + std::to_string(f) + " in bilinear regular-grid fast-path");
    M->nnz = nnz;
    M->col.resize(std::max<index_t>(nnz, 1));
    M->val.resize(std::max<index_t>(nnz, 1));
    bil_fill_kernel<<<grid_for(nd, 256), 256, 0, s>>>(g, dst.d_offsets.get(), dst.d_indices.get(), dst.d_coords.get(), dst.n_nodes(), n_dst,
                                                      M->row_ptr.get(), M->col.get(), M->val.get(), M->frac_b.get());
    fill_value_kernel<<<grid_for(ns, 256), 256, 0, s>>>(M->frac_a.get(), ns, 1.0);
    fill_value_kernel<<<grid_for(ns, 256), 256, 0, s>>>(M->area_a.get(), ns, 0.0);
    fill_value_kernel<<<grid_for(nd, 256), 256, 0, s>>>(M->area_b.get(), nd, 0.0);
    AXIS_LAUNCH_CHECK();
    M->max_row_nnz = nnz > 0 ? 4 : 0;
    cell_morton_keys(dst, M->row_key, s);
    AXIS_HIP_CHECK(hipStreamSynchronize(s));
    M->gen_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return M;
}

}  // namespace axis_hip::detail
