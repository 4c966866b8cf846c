// mesh.hip -- mesh ingest, grid detection and per-cell geometry.
//
// Reference behaviour restated here (libs/axis):
//   UnstructuredMesh ingest               topology/unstructured_mesh.hpp:41-220, mesh_factory.cpp:314-339
//   detect_regular/rectilinear_grid       detail/regular_grid_detector.hpp:64-299
//   build_node_trig_cache                 detail/trig_cache.hpp:210-267
//   get_cell_areas_spherical              src/solver/weight_generator.cpp:523-552
//   DegenerateCellHandler::classify/scan  detail/degenerate_cell_handler.hpp:115-236, 248-367
//   precompute_cap_data                   detail/spherical_cap_filter.hpp:85-251
//   compute_cell_aabbs_3d                 src/solver/weight_generator.cpp:198-262
//
// Node positions need the host libm's sin/cos bits (the clipper's sliver
// decisions depend on them), so node xyz and the trig tables are evaluated on
// the host and uploaded; everything per cell runs on the GPU.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <bit>
#include <cmath>
#include <cstring>
#include <cfloat>
#include <limits>
#include <map>
#include <mutex>

#include "axis_hip.hpp"
#include "ddtrig.hpp"
#include "geometry.hpp"
#include "host_util.hpp"
#include "morton.hpp"

namespace axis_hip {

using geo::V3;

Mesh::Mesh(const double *node_coords, index_t n_nodes, index_t ndim, const index_t *conn_offsets, index_t n_cells, const index_t *conn_indices,
           CoordinateSystem cs, const double *areas, const std::int32_t *mask)
    : n_nodes_(n_nodes), n_cells_(n_cells), cs_(cs) {
    if (n_nodes < 0 || n_cells < 0) throw InvalidArgument("Mesh: negative size");
    if (ndim != 2) throw InvalidArgument("Mesh: only [n_nodes, 2] (lon, lat) node coordinates are supported");
    if (n_cells > 0 && conn_offsets == nullptr) throw InvalidArgument("Mesh: null connectivity offsets");
    h_coords.assign(node_coords, node_coords + 2 * n_nodes);
    h_offsets.assign(conn_offsets, conn_offsets + (n_cells + 1));
    if (h_offsets.empty()) h_offsets.push_back(0);
    if (h_offsets[0] != 0) throw InvalidArgument("Mesh: conn_offsets[0] must be 0");
    for (index_t c = 0; c < n_cells; ++c) {
        index_t nv = h_offsets[c + 1] - h_offsets[c];
        if (nv < 0) throw InvalidArgument("Mesh: conn_offsets not monotone");
        max_nv_ = std::max<int>(max_nv_, static_cast<int>(std::min<index_t>(nv, 1 << 20)));
        min_nv_ = c == 0 ? static_cast<int>(std::min<index_t>(nv, 1 << 20)) : std::min<int>(min_nv_, static_cast<int>(std::min<index_t>(nv, 1 << 20)));
    }
    const index_t nidx = h_offsets[n_cells];
    h_indices.assign(conn_indices, conn_indices + nidx);
    for (index_t k = 0; k < nidx; ++k)
        if (h_indices[k] < 0 || h_indices[k] >= n_nodes) throw InvalidArgument("Mesh: node index out of range");
    if (areas) h_areas.assign(areas, areas + n_cells);
    if (mask) h_mask.assign(mask, mask + n_cells);
    hipStream_t s = nullptr;
    d_coords.resize(h_coords.size());
    d_coords.upload(h_coords.data(), h_coords.size(), s);
    d_offsets.resize(h_offsets.size());
    d_offsets.upload(h_offsets.data(), h_offsets.size(), s);
    d_indices.resize(std::max<std::size_t>(h_indices.size(), 1));
    d_indices.upload(h_indices.data(), h_indices.size(), s);
    if (areas) {
        d_areas.resize(std::max<index_t>(n_cells, 1));
        d_areas.upload(h_areas.data(), n_cells, s);
    }
    if (mask) {
        d_mask.resize(std::max<index_t>(n_cells, 1));
        d_mask.upload(h_mask.data(), n_cells, s);
    }
    AXIS_HIP_CHECK(hipStreamSynchronize(s));
}

namespace detail {

// ------------------------------------------------------------- grid axes
namespace {

__device__ inline std::uint64_t order_key(double d) {
    std::uint64_t b = static_cast<std::uint64_t>(__double_as_longlong(d));
    return (b >> 63) ? ~b : (b | 0x8000000000000000ull);
}
__device__ inline double from_order_key(std::uint64_t k) {
    std::uint64_t b = (k >> 63) ? (k & 0x7FFFFFFFFFFFFFFFull) : ~k;
    return __longlong_as_double(static_cast<long long>(b));
}

__global__ void to_keys_kernel(const double *v, index_t n, std::uint64_t *k) {
    index_t i = static_cast<index_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) k[i] = order_key(v[i]);
}

// head[i] = 1 when sorted[i] starts a new tolerance class (gap > 1e-12 to the
// previous value).
__global__ void heads_kernel(const std::uint64_t *k, index_t n, double *vals, index_t *head) {
    index_t i = static_cast<index_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double v = from_order_key(k[i]);
    vals[i] = v;
    head[i] = (i == 0 || fabs(v - from_order_key(k[i - 1])) > 1.0e-12) ? 1 : 0;
}

__global__ void compact_kernel(const double *vals, const index_t *head, const index_t *pos, index_t n, double *out) {
    index_t i = static_cast<index_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n && head[i]) out[pos[i]] = vals[i];
}

// *bad = 1 when a class spans more than 1e-12 from its head, where the
// reference's compare-with-last-kept loop could split it differently.
__global__ void span_check_kernel(const double *vals, const index_t *head, const index_t *pos, index_t n, const double *u, int *bad) {
    index_t i = static_cast<index_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n || head[i]) return;
    index_t cls = pos[i] - 1;
    if (fabs(vals[i] - u[cls]) > 1.0e-12) atomicOr(bad, 1);
}

// ---- quick pass over both coordinate columns
// Before sorting every node coordinate, two cheap looks:
//  * a lower bound on each column's tolerance-unique count: occupied bins of
//    width >= 1e-11 over [min, max] among the first kSample nodes, halved
//    (one tolerance class spans at most 1e-12, so it touches at most two
//    bins; a subset of the nodes occupies no more bins than all of them);
//    (nlon - 1) * (nlat - 1) grows with both counts, so a product of the
//    bounds above n_cells rules the grid out;
//  * each column's exact distinct values, gathered in a small hash set: the
//    reference's sort + sequential tolerance filter keeps the same values
//    when run on the distinct values alone (an exact duplicate is always
//    dropped and never becomes the last kept value).
constexpr int kSetBits = 15;            // hash slots per column
constexpr int kSetSlots = 1 << kSetBits;
constexpr int kSetMax = kSetSlots / 2;  // distinct values kept before giving up
constexpr int kBinBits = 20;            // occupancy bins per column
constexpr int kBinWords = (1 << kBinBits) / 32;
constexpr index_t kSample = 1 << 16;
constexpr int kQuickBlock = 256;

struct QuickAxes {
    unsigned long long mm[4];  // order keys: min lon, max lon, min lat, max lat
    int nonfinite;
    int n_distinct[2];
    int overflow[2];
    int occupied[2];
};

__global__ void __launch_bounds__(kQuickBlock) minmax_kernel(const double *lon, const double *lat, index_t n, QuickAxes *q) {
    std::uint64_t r[4] = {~0ull, 0ull, ~0ull, 0ull};
    bool bad = false;
    for (index_t i = static_cast<index_t>(blockIdx.x) * kQuickBlock + threadIdx.x; i < n; i += static_cast<index_t>(gridDim.x) * kQuickBlock) {
        const double a = lon[i], b = lat[i];
        bad = bad || !isfinite(a) || !isfinite(b);
        const std::uint64_t ka = order_key(a), kb = order_key(b);
        r[0] = min(r[0], ka);
        r[1] = max(r[1], ka);
        r[2] = min(r[2], kb);
        r[3] = max(r[3], kb);
    }
    for (int o = 32; o > 0; o >>= 1)
        for (int k = 0; k < 4; ++k) {
            const std::uint64_t v = static_cast<std::uint64_t>(__shfl_xor(static_cast<unsigned long long>(r[k]), o));
            r[k] = (k & 1) ? max(r[k], v) : min(r[k], v);
        }
    __shared__ unsigned long long part[kQuickBlock / 64][4];
    __shared__ int any_bad;
    if (threadIdx.x == 0) any_bad = 0;
    __syncthreads();
    if (bad) any_bad = 1;
    if ((threadIdx.x & 63) == 0)
        for (int k = 0; k < 4; ++k) part[threadIdx.x / 64][k] = r[k];
    __syncthreads();
    if (threadIdx.x < 4) {
        const int k = threadIdx.x;
        std::uint64_t v = part[0][k];
        for (int w = 1; w < kQuickBlock / 64; ++w) v = (k & 1) ? max(v, static_cast<std::uint64_t>(part[w][k])) : min(v, static_cast<std::uint64_t>(part[w][k]));
        if (k & 1)
            atomicMax(&q->mm[k], static_cast<unsigned long long>(v));
        else
            atomicMin(&q->mm[k], static_cast<unsigned long long>(v));
        if (k == 0 && any_bad) atomicOr(&q->nonfinite, 1);
    }
}

__device__ inline std::uint32_t mix_key(std::uint64_t k) {
    k ^= k >> 33;
    k *= 0xff51afd7ed558ccdull;
    k ^= k >> 33;
    k *= 0xc4ceb9fe1a85ec53ull;
    k ^= k >> 33;
    return static_cast<std::uint32_t>(k);
}

// occupancy bits of the first n nodes (finite inputs)
__global__ void __launch_bounds__(kQuickBlock) sample_bins_kernel(const double *lon, const double *lat, index_t n, const QuickAxes *q,
                                                                  std::uint32_t *bins) {
    const index_t i = static_cast<index_t>(blockIdx.x) * kQuickBlock + threadIdx.x;
    if (i >= n) return;
    for (int c = 0; c < 2; ++c) {
        const double v = c == 0 ? lon[i] : lat[i];
        const double mn = from_order_key(q->mm[2 * c]), mx = from_order_key(q->mm[2 * c + 1]);
        const double range = mx - mn;
        const double nb = fmax(1.0, fmin(static_cast<double>(1 << kBinBits), floor(range / 1.0e-11)));
        const double x = range > 0.0 ? (v - mn) * (nb / range) : 0.0;
        const int bin = static_cast<int>(fmin(nb - 1.0, fmax(0.0, floor(x))));
        std::uint32_t *bw = bins + c * kBinWords + (bin >> 5);
        const std::uint32_t bit = 1u << (bin & 31);
        if (!(__hip_atomic_load(bw, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) & bit)) atomicOr(bw, bit);
    }
}

// occupied-bin counts: each thread a run of words, one atomic per workgroup
// (a workgroup's words lie in one column's half)
constexpr int kCountWords = 16;  // words per thread
__global__ void __launch_bounds__(256) count_bins_kernel(const std::uint32_t *bins, QuickAxes *q) {
    const int w0 = (static_cast<int>(blockIdx.x) * 256 + static_cast<int>(threadIdx.x)) * kCountWords;  // < 2 * kBinWords
    int c = 0;
    for (int k = 0; k < kCountWords; ++k) c += __popc(bins[w0 + k]);
    for (int o = 32; o > 0; o >>= 1) c += __shfl_xor(c, o);
    __shared__ int part[4];
    if ((threadIdx.x & 63) == 0) part[threadIdx.x / 64] = c;
    __syncthreads();
    if (threadIdx.x == 0) atomicAdd(&q->occupied[(static_cast<int>(blockIdx.x) * 256 * kCountWords) / kBinWords], part[0] + part[1] + part[2] + part[3]);
}

// distinct values of both columns (finite inputs: order keys are then never
// 0, the empty-slot marker); a lane skips a value equal to its left
// neighbour's, which removes most repeats of structured node orders
__global__ void __launch_bounds__(kQuickBlock) distinct_kernel(const double *lon, const double *lat, index_t n, QuickAxes *q,
                                                               unsigned long long *set) {
    const index_t i = static_cast<index_t>(blockIdx.x) * kQuickBlock + threadIdx.x;
    const bool act = i < n;
    for (int c = 0; c < 2; ++c) {
        const double v = act ? (c == 0 ? lon[i] : lat[i]) : 0.0;
        const unsigned long long key = act ? order_key(v) : 0ull;
        const unsigned long long left = __shfl_up(key, 1);
        if (!act || ((threadIdx.x & 63) != 0 && left == key)) continue;
        if (!isfinite(v)) {  // reported to the host, which then takes the sorting path
            atomicOr(&q->nonfinite, 1);
            continue;
        }
        if (__hip_atomic_load(&q->overflow[c], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) continue;
        unsigned long long *tab = set + c * kSetSlots;
        std::uint32_t h = mix_key(key) & (kSetSlots - 1);
        for (int probe = 0; probe < kSetSlots; ++probe) {
            unsigned long long cur = __hip_atomic_load(&tab[h], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (cur == 0) {
                cur = atomicCAS(&tab[h], 0ull, key);
                if (cur == 0) {
                    if (atomicAdd(&q->n_distinct[c], 1) + 1 > kSetMax) atomicOr(&q->overflow[c], 1);
                    break;
                }
            }
            if (cur == key) break;
            if (__hip_atomic_load(&q->overflow[c], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) break;
            h = (h + 1) & (kSetSlots - 1);
        }
    }
}

__global__ void gather_set_kernel(const unsigned long long *set, double *out, int *pos) {
    const int t = static_cast<int>(blockIdx.x) * blockDim.x + threadIdx.x;  // < 2 * kSetSlots
    const unsigned long long k = set[t];
    if (k == 0) return;
    const int p = atomicAdd(&pos[t / kSetSlots], 1);
    if (p < kSetMax) out[(t / kSetSlots) * kSetMax + p] = from_order_key(k);  // an overflowed set holds more
}

std::vector<double> host_unique(std::vector<double> s) {
    std::sort(s.begin(), s.end());
    std::vector<double> u;
    if (s.empty()) return u;
    u.push_back(s[0]);
    for (std::size_t i = 1; i < s.size(); ++i)
        if (std::abs(s[i] - u.back()) > 1.0e-12) u.push_back(s[i]);
    return u;
}

// Tolerance-unique sorted values of one coordinate column (device).  Returns
// the count; fills `out` only when `want` says so.
index_t unique_column(const double *d_col, const double *h_col, index_t n, bool (*want)(index_t, void *), void *ctx, std::vector<double> &out,
                      hipStream_t s) {
    out.clear();
    if (n == 0) return 0;
    DeviceBuffer<std::uint64_t> keys(n);
    to_keys_kernel<<<grid_for(n, 256), 256, 0, s>>>(d_col, n, keys.get());
    AXIS_LAUNCH_CHECK();
    radix_sort_keys_u64(keys.get(), static_cast<std::size_t>(n), s);
    DeviceBuffer<double> vals(n);
    DeviceBuffer<index_t> head(n), pos(n + 1);
    DeviceBuffer<int> bad(1);
    AXIS_HIP_CHECK(hipMemsetAsync(bad.get(), 0, sizeof(int), s));
    heads_kernel<<<grid_for(n, 256), 256, 0, s>>>(keys.get(), n, vals.get(), head.get());
    AXIS_LAUNCH_CHECK();
    exclusive_scan_i64(head.get(), pos.get(), n, s);
    index_t count = 0;
    read_back(&count, pos.get() + n, 1, s);
    DeviceBuffer<double> u(count);
    compact_kernel<<<grid_for(n, 256), 256, 0, s>>>(vals.get(), head.get(), pos.get(), n, u.get());
    AXIS_LAUNCH_CHECK();
    span_check_kernel<<<grid_for(n, 256), 256, 0, s>>>(vals.get(), head.get(), pos.get(), n, u.get(), bad.get());
    AXIS_LAUNCH_CHECK();
    int h_bad = 0;
    read_back(&h_bad, bad.get(), 1, s);
    if (h_bad) {
        // a tolerance class wider than 1e-12: replay the reference's sequential filter exactly
        out = host_unique(std::vector<double>(h_col, h_col + n));
        return static_cast<index_t>(out.size());
    }
    if (want(count, ctx)) {
        out.resize(count);
        u.download(out.data(), count, s);
        AXIS_HIP_CHECK(hipStreamSynchronize(s));
    }
    return count;
}

}  // namespace

// Whether node p holds the longitude of node p % nlon and the latitude of
// node (p / nlon) nlon, bit for bit, and finite values (a longitude-fastest
// nlon x nlat node grid); one flag per failing wave.
__global__ void __launch_bounds__(256) canonical_grid_kernel(const double *__restrict__ lon, const double *__restrict__ lat, index_t n, index_t nlon,
                                                             int *__restrict__ bad) {
    const index_t p = static_cast<index_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    bool ok = true;
    if (p < n) {
        const double a = lon[p], b = lat[p];
        ok = isfinite(a) && isfinite(b) && __double_as_longlong(a) == __double_as_longlong(lon[p % nlon]) &&
             __double_as_longlong(b) == __double_as_longlong(lat[(p / nlon) * nlon]);
    }
    if (__ballot(!ok) != 0ull && (threadIdx.x & 63) == 0) atomicOr(bad, 1);
}

// total order for the distinct coordinate values: by value, ties (-0, +0) by bits
inline bool value_then_bits(double a, double b) {
    if (a < b) return true;
    if (b < a) return false;
    std::uint64_t x, y;
    std::memcpy(&x, &a, sizeof x);
    std::memcpy(&y, &b, sizeof y);
    return x < y;
}

// Node and cell counts of some nlon x nlat grid (nlon, nlat >= 2):
// nlon + nlat = nodes - cells + 1 = S and nlon nlat = nodes, so
// (nlon, nlat) = (S +- d) / 2 with d^2 = S^2 - 4 nodes.
bool grid_counts(const Mesh &m, index_t *na = nullptr, index_t *nb = nullptr) {
    if (m.n_cells() == 0) return false;
    const __int128 S = static_cast<__int128>(m.n_nodes()) - m.n_cells() + 1, P = m.n_nodes();
    const __int128 disc = S * S - 4 * P;
    if (S < 4 || disc < 0) return false;
    const index_t r = static_cast<index_t>(std::llround(std::sqrt(static_cast<long double>(disc))));
    for (index_t d = r - 1; d <= r + 1; ++d)
        if (d >= 0 && static_cast<__int128>(d) * d == disc && (S + d) % 2 == 0 && (S - d) / 2 >= 2) {
            if (na) *na = static_cast<index_t>((S + d) / 2);
            if (nb) *nb = static_cast<index_t>((S - d) / 2);
            return true;
        }
    return false;
}

// Exact distinct coordinate values of a mesh whose nodes form a
// longitude-fastest grid (checked on the device), read from the first row
// and column of the host coordinates; false when the nodes are in another
// layout.
bool canonical_axes(const Mesh &m, index_t na, index_t nb, std::vector<double> &xlon, std::vector<double> &xlat, hipStream_t s) {
    const index_t n = m.n_nodes();
    for (int t = 0; t < 2; ++t) {
        const index_t nlon = t == 0 ? na : nb, nlat = t == 0 ? nb : na;
        if (t == 1 && na == nb) continue;  // (the same layout again)
        DeviceBuffer<int> bad(1);
        AXIS_HIP_CHECK(hipMemsetAsync(bad.get(), 0, sizeof(int), s));
        canonical_grid_kernel<<<grid_for(n, 256), 256, 0, s>>>(m.d_coords.get(), m.d_coords.get() + n, n, nlon, bad.get());
        AXIS_LAUNCH_CHECK();
        int h_bad = 1;
        read_back(&h_bad, bad.get(), 1, s);
        if (h_bad) continue;
        auto distinct = [](std::vector<double> v) {
            std::sort(v.begin(), v.end(), value_then_bits);
            v.erase(std::unique(v.begin(), v.end(), [](double a, double b) { return !value_then_bits(a, b) && !value_then_bits(b, a); }), v.end());
            return v;
        };
        const double *lon = m.h_coords.data(), *lat = m.h_coords.data() + n;
        std::vector<double> vlat(nlat);
        for (index_t j = 0; j < nlat; ++j) vlat[j] = lat[j * nlon];
        xlon = distinct(std::vector<double>(lon, lon + nlon));
        xlat = distinct(std::move(vlat));
        return true;
    }
    return false;
}

GridAxes unique_axes(const Mesh &m, hipStream_t s) {
    GridAxes ax;
    const index_t n_cells = m.n_cells();
    if (n_cells == 0) return ax;
    ax.all_quads = m.min_cell_vertices() == 4 && m.max_cell_vertices() == 4;
    if (!ax.all_quads) return ax;
    const index_t n = m.n_nodes();
    PhaseTrace tr("axes", s);
    if (n > 0) {
        // quick pass: refute with the lower bounds, or take both columns'
        // distinct values from the hash sets.  Node and cell counts of some
        // nlon x nlat grid leave nothing for the bounds to refute: then the
        // distinct pass alone runs (and checks for non-finite values itself).
        DeviceBuffer<QuickAxes> dq(1);
        QuickAxes hq{};
        hq.mm[0] = hq.mm[2] = ~0ull;
        const double *lon = m.d_coords.get(), *lat = m.d_coords.get() + n;
        AXIS_HIP_CHECK(hipMemcpyAsync(dq.get(), &hq, sizeof hq, hipMemcpyHostToDevice, s));
        bool run_distinct = true;
        index_t na = 0, nb = 0;
        const bool counts = grid_counts(m, &na, &nb);
        std::vector<double> xlon, xlat;
        if (counts && canonical_axes(m, na, nb, xlon, xlat, s)) {
            std::vector<double> ulon = host_unique(xlon), ulat = host_unique(xlat);
            const index_t nlon = static_cast<index_t>(ulon.size()), nlat = static_cast<index_t>(ulat.size());
            if (nlon >= 2 && nlat >= 2 && (nlon - 1) * (nlat - 1) == n_cells) {
                ax.ulon = std::move(ulon);
                ax.ulat = std::move(ulat);
                ax.xlon = std::move(xlon);
                ax.xlat = std::move(xlat);
            }
            tr.mark("canonical");
            return ax;
        }
        if (!counts) {
            DeviceBuffer<std::uint32_t> bins(2 * kBinWords);
            AXIS_HIP_CHECK(hipMemsetAsync(bins.get(), 0, sizeof(std::uint32_t) * 2 * kBinWords, s));
            // (256 workgroups: their closing atomics on four words serialise)
            const unsigned mm_grid = static_cast<unsigned>(std::min<index_t>((n + kQuickBlock - 1) / kQuickBlock, 256));
            minmax_kernel<<<mm_grid, kQuickBlock, 0, s>>>(lon, lat, n, dq.get());
            AXIS_LAUNCH_CHECK();
            const index_t ns = std::min(n, kSample);
            sample_bins_kernel<<<grid_for(ns, kQuickBlock), kQuickBlock, 0, s>>>(lon, lat, ns, dq.get(), bins.get());
            AXIS_LAUNCH_CHECK();
            count_bins_kernel<<<2 * kBinWords / (256 * kCountWords), 256, 0, s>>>(bins.get(), dq.get());
            AXIS_LAUNCH_CHECK();
            read_back(&hq, dq.get(), 1, s);
            tr.mark("bound");
            if (hq.nonfinite) {
                run_distinct = false;  // (the bins are meaningless with non-finite coordinates)
            } else {
                const index_t lb_lon = (hq.occupied[0] + 1) / 2, lb_lat = (hq.occupied[1] + 1) / 2;
                if (lb_lon >= 2 && lb_lat >= 2 && (lb_lon - 1) > n_cells / (lb_lat - 1)) return ax;
            }
        }
        if (run_distinct) {
            DeviceBuffer<unsigned long long> set(2 * kSetSlots);
            DeviceBuffer<double> vals(2 * kSetMax);
            DeviceBuffer<int> pos(2);
            AXIS_HIP_CHECK(hipMemsetAsync(set.get(), 0, sizeof(unsigned long long) * 2 * kSetSlots, s));
            AXIS_HIP_CHECK(hipMemsetAsync(pos.get(), 0, 2 * sizeof(int), s));
            distinct_kernel<<<grid_for(n, kQuickBlock), kQuickBlock, 0, s>>>(lon, lat, n, dq.get(), set.get());
            AXIS_LAUNCH_CHECK();
            gather_set_kernel<<<2 * kSetSlots / 256, 256, 0, s>>>(set.get(), vals.get(), pos.get());
            AXIS_LAUNCH_CHECK();
            read_back(&hq, dq.get(), 1, s);
            if (!hq.nonfinite && !hq.overflow[0] && !hq.overflow[1]) {
                xlon.assign(hq.n_distinct[0], 0.0);
                xlat.assign(hq.n_distinct[1], 0.0);
                if (!xlon.empty())
                    AXIS_HIP_CHECK(hipMemcpyAsync(xlon.data(), vals.get(), sizeof(double) * xlon.size(), hipMemcpyDeviceToHost, s));
                if (!xlat.empty())
                    AXIS_HIP_CHECK(hipMemcpyAsync(xlat.data(), vals.get() + kSetMax, sizeof(double) * xlat.size(), hipMemcpyDeviceToHost, s));
                AXIS_HIP_CHECK(hipStreamSynchronize(s));
                std::sort(xlon.begin(), xlon.end(), value_then_bits);
                std::sort(xlat.begin(), xlat.end(), value_then_bits);
                std::vector<double> ulon = host_unique(xlon), ulat = host_unique(xlat);
                const index_t nlon = static_cast<index_t>(ulon.size()), nlat = static_cast<index_t>(ulat.size());
                if (nlon >= 2 && nlat >= 2 && (nlon - 1) * (nlat - 1) == n_cells) {
                    ax.ulon = std::move(ulon);
                    ax.ulat = std::move(ulat);
                    ax.xlon = std::move(xlon);
                    ax.xlat = std::move(xlat);
                }
                tr.mark("distinct");
                return ax;
            }
        }
    }
    // the lat count is needed to decide whether the lon values are worth downloading
    struct Ctx {
        index_t n_cells, other;
    } ctx{n_cells, 0};
    auto want_lat = [](index_t, void *) { return false; };
    std::vector<double> tmp;
    index_t nlat = unique_column(m.d_coords.get() + n, m.h_coords.data() + n, n, want_lat, nullptr, tmp, s);
    ctx.other = nlat;
    // (nlon - 1) * (nlat - 1) == n_cells needs nlat - 1 to divide n_cells:
    // otherwise no lon count can make the mesh rectilinear, skip that sort
    if (nlat < 2 || n_cells % (nlat - 1) != 0) return ax;
    auto want_lon = [](index_t cnt, void *p) {
        auto *c = static_cast<Ctx *>(p);
        return cnt >= 2 && c->other >= 2 && (cnt - 1) * (c->other - 1) == c->n_cells;
    };
    index_t nlon = unique_column(m.d_coords.get(), m.h_coords.data(), n, want_lon, &ctx, ax.ulon, s);
    if (nlon >= 2 && nlat >= 2 && (nlon - 1) * (nlat - 1) == n_cells) {
        if (tmp.empty()) {
            auto want_all = [](index_t, void *) { return true; };
            unique_column(m.d_coords.get() + n, m.h_coords.data() + n, n, want_all, nullptr, ax.ulat, s);
        } else {
            ax.ulat = tmp;
        }
        if (ax.ulon.empty()) {
            auto want_all = [](index_t, void *) { return true; };
            unique_column(m.d_coords.get(), m.h_coords.data(), n, want_all, nullptr, ax.ulon, s);
        }
    } else {
        ax.ulon.clear();
        ax.ulat.clear();
    }
    return ax;
}

bool rectilinear_from_axes(const Mesh &m, const GridAxes &ax) {
    if (!ax.all_quads || ax.ulon.size() < 2 || ax.ulat.size() < 2) return false;
    return static_cast<index_t>(ax.ulon.size() - 1) * static_cast<index_t>(ax.ulat.size() - 1) == m.n_cells();
}

bool regular_from_axes(const Mesh &m, const GridAxes &ax, MeshGeom &g) {
    g.regular = false;
    if (!rectilinear_from_axes(m, ax)) return false;
    auto deltas = [](const std::vector<double> &u) {
        std::vector<double> d(u.size() - 1);
        for (std::size_t i = 0; i + 1 < u.size(); ++i) d[i] = u[i + 1] - u[i];
        return d;
    };
    auto uniform = [](const std::vector<double> &d) {
        if (d.empty()) return false;
        double mn = d[0], mx = d[0], sum = 0.0;
        for (double x : d) {
            mn = std::min(mn, x);
            mx = std::max(mx, x);
            sum += x;
        }
        if (mn <= 0.0) return false;
        double mean = sum / static_cast<double>(d.size());
        return (mx - mn) < 1.0e-10 * mean;
    };
    auto dl = deltas(ax.ulon), da = deltas(ax.ulat);
    if (!uniform(dl) || !uniform(da)) return false;
    double sl = 0.0, sa = 0.0;
    for (double x : dl) sl += x;
    for (double x : da) sa += x;
    g.regular = true;
    g.ni = static_cast<index_t>(dl.size());
    g.nj = static_cast<index_t>(da.size());
    g.lon_min = ax.ulon.front();
    g.delta_lon = sl / static_cast<double>(g.ni);
    g.lat_min = ax.ulat.front();
    g.delta_lat = sa / static_cast<double>(g.nj);
    return true;
}

// ----------------------------------------------------------- per-cell pass
namespace {

constexpr int kCellBlock = 256;

__device__ inline V3 node(const double *xyz, index_t n) { return V3{xyz[3 * n], xyz[3 * n + 1], xyz[3 * n + 2]}; }

// A cell's vertices: through its node indices, or held in registers (a
// quad's four, loaded together up front).
struct NodeVtx {
    const double *xyz;
    const index_t *idx;
    __device__ V3 operator()(int k) const { return node(xyz, idx[k]); }
};
struct QuadVtx {
    V3 v[4];
    __device__ V3 operator()(int k) const { return v[k]; }
};

// DegenerateCellHandler::classify on the first min(nv, 32) vertices.
// Returns the classification and the fan area it used.
template <class Vtx>
__device__ int classify_cell(const Vtx &P, int nv, double &fan_area) {
    fan_area = 0.0;
    if (nv < 3) return 1;
    // area (spherical_clipper.hpp:112-137)
    double total = 0.0;
    V3 a = P(0);
    V3 b = P(1);
    for (int i = 1; i < nv - 1; ++i) {
        V3 c = P(i + 1);
        double num;
        total += geo::fan_term(a, b, c, num);
        b = c;
    }
    fan_area = fabs(total);
    if (fan_area < 1e-25) return 1;
    // count_unique_vertices (threshold 1e-15)
    const double t2 = 1e-15 * 1e-15;
    int uniq = 1;
    for (int i = 1; i < nv; ++i) {
        V3 vi = P(i);
        bool dup = false;
        for (int j = 0; j < i; ++j) {
            V3 vj = P(j);
            V3 d{vi.x - vj.x, vi.y - vj.y, vi.z - vj.z};
            if (geo::length_sq(d) < t2) {
                dup = true;
                break;
            }
        }
        if (!dup) ++uniq;
    }
    if (uniq < 3) return 2;
    // is_self_intersecting
    if (nv >= 4) {
        for (int i = 0; i < nv; ++i) {
            int in = (i + 1) % nv;
            V3 pa = P(i), pb = P(in);
            V3 nab = geo::cross(pa, pb);
            for (int j = i + 2; j < nv; ++j) {
                int jn = (j + 1) % nv;
                if (jn == i) continue;
                V3 pc = P(j), pd = P(jn);
                double dc = geo::dot(nab, pc), dd = geo::dot(nab, pd);
                if (dc * dd >= 0.0) continue;
                V3 ncd = geo::cross(pc, pd);
                double da = geo::dot(ncd, pa), db = geo::dot(ncd, pb);
                if (da * db >= 0.0) continue;
                return 3;
            }
        }
    }
    return 0;
}

__device__ inline float f_down(double v) { return nextafterf(__double2float_rd(v), -INFINITY); }
__device__ inline float f_up(double v) { return nextafterf(__double2float_ru(v), INFINITY); }

// one cell's flags, area, cap and box; A gives its vertices' xyz, Dg those
// of the degree reading of its coordinates (the same when `same`); NV, when
// not 0, is the cell's vertex count known at compile time (all loops then
// unroll over registers)
template <int NV, class VA, class VD>
__device__ inline void cell_prep_one(index_t c, int nv_dyn, index_t nv_all_dyn, const VA &A, const VD &Dg, bool same, const double *mesh_areas,
                                     const std::int32_t *mask, double *area_out, std::int32_t *flags_out, double *cap_out, float *box_out,
                                     index_t *recheck, int *n_recheck) {
    const int nv = NV ? NV : nv_dyn;
    const index_t nv_all = NV ? NV : nv_all_dyn;
    // degenerate scan (always on the degree interpretation of the coordinates)
    double fan_deg;
    int cls = classify_cell(Dg, nv, fan_deg);
    // cell area: mesh areas when given, else the fan of the polygon
    double area;
    if (mesh_areas) {
        area = mesh_areas[c];
    } else if (same) {
        area = fan_deg;
    } else {
        double total = 0.0;
        if (nv >= 3) {
            V3 a = A(0);
            V3 bb = A(1);
            for (int i = 1; i < nv - 1; ++i) {
                V3 cc = A(i + 1);
                double num;
                total += geo::fan_term(a, bb, cc, num);
                bb = cc;
            }
        }
        area = fabs(total);
    }
    int flags = (cls != 0 ? 1 : 0) | ((mask && mask[c] == 0) ? 2 : 0) | (area <= 0.0 ? 4 : 0);
    // decisions taken on atan2 bits near zero are replayed on the host
    if ((nv >= 3 && fan_deg < 1e-20) || (!mesh_areas && area < 1e-20)) {
        int k = atomicAdd(n_recheck, 1);
        recheck[k] = c;
    }
    area_out[c] = area;
    flags_out[c] = flags;
    // cap data (spherical_cap_filter.hpp:85-251) over all vertices
    V3 cen{0.0, 0.0, 0.0};
    double rad = 0.0;
    if (nv_all >= 1) {
        double sx = 0, sy = 0, sz = 0;
        for (index_t k = 0; k < nv_all; ++k) {
            V3 p = A(static_cast<int>(k));
            sx += p.x;
            sy += p.y;
            sz += p.z;
        }
        double inv = 1.0 / static_cast<double>(nv_all);
        cen = geo::normalize(V3{sx * inv, sy * inv, sz * inv});
        for (index_t k = 0; k < nv_all; ++k) {
            double d = geo::dot(cen, A(static_cast<int>(k)));
            d = fmin(1.0, fmax(-1.0, d));
            double ang = acos(d);
            if (ang > rad) rad = ang;
        }
    }
    cap_out[4 * c + 0] = cen.x;
    cap_out[4 * c + 1] = cen.y;
    cap_out[4 * c + 2] = cen.z;
    cap_out[4 * c + 3] = rad;
    // float AABB with the reference's cell-size-scaled dilation, rounded
    // outward one extra float ulp so it contains the reference's box
    double mn[3] = {DBL_MAX, DBL_MAX, DBL_MAX}, mx[3] = {-DBL_MAX, -DBL_MAX, -DBL_MAX};
    for (index_t k = 0; k < nv_all; ++k) {
        V3 p = A(static_cast<int>(k));
        mn[0] = fmin(mn[0], p.x);
        mn[1] = fmin(mn[1], p.y);
        mn[2] = fmin(mn[2], p.z);
        mx[0] = fmax(mx[0], p.x);
        mx[1] = fmax(mx[1], p.y);
        mx[2] = fmax(mx[2], p.z);
    }
    double ext = fmax(fmax(mx[0] - mn[0], mx[1] - mn[1]), mx[2] - mn[2]);
    double eps = fmin(fmax(0.25 * ext, 1.0e-6), 0.02);
    for (int a = 0; a < 3; ++a) {
        box_out[6 * c + a] = nv_all ? f_down(mn[a] - eps) : INFINITY;
        box_out[6 * c + 3 + a] = nv_all ? f_up(mx[a] + eps) : -INFINITY;
    }
}

__global__ void __launch_bounds__(kCellBlock) cell_prep_kernel(index_t n_cells, const index_t *off, const index_t *ind, const double *xyz,
                                                               const double *xyz_deg, const double *mesh_areas, const std::int32_t *mask,
                                                               double *area_out, std::int32_t *flags_out, double *cap_out, float *box_out,
                                                               index_t *recheck, int *n_recheck) {
    index_t c = static_cast<index_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (c >= n_cells) return;
    const index_t b = off[c];
    const index_t nv_all = off[c + 1] - b;
    const int nv = static_cast<int>(nv_all < 32 ? nv_all : 32);
    const index_t *idx = ind + b;
    const bool same = xyz_deg == xyz;
    if (same && nv_all == 4) {
        // a quad: its four vertices loaded together, once
        QuadVtx Q;
        const index_t i0 = idx[0], i1 = idx[1], i2 = idx[2], i3 = idx[3];
        Q.v[0] = node(xyz, i0);
        Q.v[1] = node(xyz, i1);
        Q.v[2] = node(xyz, i2);
        Q.v[3] = node(xyz, i3);
        cell_prep_one<4>(c, nv, nv_all, Q, Q, true, mesh_areas, mask, area_out, flags_out, cap_out, box_out, recheck, n_recheck);
    } else {
        cell_prep_one<0>(c, nv, nv_all, NodeVtx{xyz, idx}, NodeVtx{xyz_deg, idx}, same, mesh_areas, mask, area_out, flags_out, cap_out, box_out, recheck,
                      n_recheck);
    }
}

// host replay of the degenerate classification for the rechecked cells
int host_classify(const double *xyz, const index_t *idx, int nv, double &fan_area) {
    using geo::V3;
    auto P = [&](int i) { return V3{xyz[3 * idx[i]], xyz[3 * idx[i] + 1], xyz[3 * idx[i] + 2]}; };
    fan_area = 0.0;
    if (nv < 3) return 1;
    double total = 0.0;
    for (int i = 1; i < nv - 1; ++i) {
        double num;
        total += geo::fan_term(P(0), P(i), P(i + 1), num);
    }
    fan_area = std::fabs(total);
    if (fan_area < 1e-25) return 1;
    const double t2 = 1e-15 * 1e-15;
    int uniq = 1;
    for (int i = 1; i < nv; ++i) {
        bool dup = false;
        for (int j = 0; j < i; ++j) {
            V3 a = P(i), b = P(j);
            V3 d{a.x - b.x, a.y - b.y, a.z - b.z};
            if (geo::length_sq(d) < t2) {
                dup = true;
                break;
            }
        }
        if (!dup) ++uniq;
    }
    if (uniq < 3) return 2;
    if (nv >= 4) {
        for (int i = 0; i < nv; ++i) {
            int in = (i + 1) % nv;
            for (int j = i + 2; j < nv; ++j) {
                int jn = (j + 1) % nv;
                if (jn == i) continue;
                V3 nab = geo::cross(P(i), P(in));
                double dc = geo::dot(nab, P(j)), dd = geo::dot(nab, P(jn));
                if (dc * dd >= 0.0) continue;
                V3 ncd = geo::cross(P(j), P(jn));
                double da = geo::dot(ncd, P(i)), db = geo::dot(ncd, P(in));
                if (da * db >= 0.0) continue;
                return 3;
            }
        }
    }
    return 0;
}

__device__ inline int lower_bound_dev(const double *u, int n, double x) {
    int lo = 0, hi = n;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (u[mid] < x)
            lo = mid + 1;
        else
            hi = mid;
    }
    return lo;
}

// xyz of every node of a rectilinear degree mesh from the sin/cos tables of
// its axis values: tab = [ulon | ulat | sin lon | cos lon | sin lat | cos lat].
// A node whose lon or lat equals no axis value bit for bit is listed for the
// host (its xyz is left zero here).
__global__ void axis_xyz_kernel(const double *__restrict__ lon, const double *__restrict__ lat, index_t n, const double *__restrict__ tab, int nlon,
                                int nlat, double *__restrict__ xyz, index_t *__restrict__ fix, int *__restrict__ n_fix, int fix_cap) {
    const index_t i = static_cast<index_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double *ulon = tab, *ulat = tab + nlon, *slon = ulat + nlat, *clon = slon + nlon, *slat = clon + nlon, *clat = slat + nlat;
    const double x = lon[i], y = lat[i];
    const int a = lower_bound_dev(ulon, nlon, x), b = lower_bound_dev(ulat, nlat, y);
    const bool ok = a < nlon && b < nlat && __double_as_longlong(ulon[a]) == __double_as_longlong(x) &&
                    __double_as_longlong(ulat[b]) == __double_as_longlong(y);
    double px = 0.0, py = 0.0, pz = 0.0;
    if (ok) {
        const double cl = clat[b];
        px = cl * clon[a];
        py = cl * slon[a];
        pz = slat[b];
    } else {
        const int k = atomicAdd(n_fix, 1);
        if (k < fix_cap) fix[k] = i;
    }
    xyz[3 * i] = px;
    xyz[3 * i + 1] = py;
    xyz[3 * i + 2] = pz;
}

__global__ void scatter_xyz_kernel(const index_t *__restrict__ nodes, const double *__restrict__ vals, int n, double *__restrict__ xyz) {
    const int k = static_cast<int>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (k >= n) return;
    for (int d = 0; d < 3; ++d) xyz[3 * nodes[k] + d] = vals[3 * k + d];
}

constexpr int kXyzFixCap = 8192;

// xyz of every node of a degree mesh from double-double sin/cos
// (lonlat_to_xyz, spherical_geometry.hpp:80-83: x = cos lat cos lon,
// y = cos lat sin lon, z = sin lat, one product each).  Bit i of word i/64 of
// flag_lat / flag_lon marks a node whose latitude / longitude has a sine or
// cosine within dd::kLibmBand ULP of a rounding midpoint: its xyz here may be
// one ULP off host libm's until trig_fix_kernel.  Nodes outside the
// reduction's range, or with a sine or cosine below dd::kTinyResult, are
// listed in `full` for the host.
__global__ void __launch_bounds__(256) trig_xyz_kernel(const double *__restrict__ lon, const double *__restrict__ lat, index_t n, index_t words,
                                                       double *__restrict__ xyz, std::uint64_t *__restrict__ flag_lat,
                                                       std::uint64_t *__restrict__ flag_lon, index_t *__restrict__ full, int *__restrict__ n_full,
                                                       int full_cap) {
    const index_t i = static_cast<index_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    bool f_lat = false, f_lon = false;
    if (i < n) {
        dd::DD sl, cl, so, co;
        if (dd::sincos_dd(lat[i] * kDeg2Rad, sl, cl) && dd::sincos_dd(lon[i] * kDeg2Rad, so, co) && !dd::tiny(sl) && !dd::tiny(cl) &&
            !dd::tiny(so) && !dd::tiny(co)) {
            f_lat = dd::near_midpoint(sl) || dd::near_midpoint(cl);
            f_lon = dd::near_midpoint(so) || dd::near_midpoint(co);
            xyz[3 * i] = cl.hi * co.hi;
            xyz[3 * i + 1] = cl.hi * so.hi;
            xyz[3 * i + 2] = sl.hi;
        } else {
            const int k = atomicAdd(n_full, 1);
            if (k < full_cap) full[k] = i;
        }
    }
    // one word per wave (the block is four whole waves)
    const std::uint64_t bl = __ballot(f_lat), bo = __ballot(f_lon);
    const index_t w = i >> 6;
    if ((threadIdx.x & 63) == 0 && w < words) {
        flag_lat[w] = bl;
        flag_lon[w] = bo;
    }
}

// The flagged nodes again, each sine and cosine of a flagged angle replaced by
// host libm's double (libm_bits: the last bit of host libm's sin lat, cos lat,
// sin lon, cos lon, words [0, words), [words, 2 words), ...).
__global__ void __launch_bounds__(256) trig_fix_kernel(const double *__restrict__ lon, const double *__restrict__ lat, index_t n, index_t words,
                                                       const std::uint64_t *__restrict__ flag_lat, const std::uint64_t *__restrict__ flag_lon,
                                                       const std::uint64_t *__restrict__ libm_bits, double *__restrict__ xyz) {
    const index_t i = static_cast<index_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const index_t w = i >> 6;
    const int t = static_cast<int>(i & 63);
    const bool fl = (flag_lat[w] >> t) & 1, fo = (flag_lon[w] >> t) & 1;
    if (!fl && !fo) return;
    dd::DD sl, cl, so, co;
    dd::sincos_dd(lat[i] * kDeg2Rad, sl, cl);  // (in range: the node was not listed for the host)
    dd::sincos_dd(lon[i] * kDeg2Rad, so, co);
    auto bit = [&](int q) { return static_cast<unsigned>((libm_bits[q * words + w] >> t) & 1); };
    const double vsl = fl ? dd::pick(sl, bit(0)) : sl.hi, vcl = fl ? dd::pick(cl, bit(1)) : cl.hi;
    const double vso = fo ? dd::pick(so, bit(2)) : so.hi, vco = fo ? dd::pick(co, bit(3)) : co.hi;
    xyz[3 * i] = vcl * vco;
    xyz[3 * i + 1] = vcl * vso;
    xyz[3 * i + 2] = vsl;
}

__global__ void scatter_recheck_kernel(const index_t *__restrict__ cells, const double *__restrict__ area, const std::int32_t *__restrict__ flags, int n,
                                       double *__restrict__ d_area, std::int32_t *__restrict__ d_flags) {
    const int k = static_cast<int>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (k >= n) return;
    d_area[cells[k]] = area[k];
    d_flags[cells[k]] = flags[k];
}

}  // namespace

void node_xyz_one(const Mesh &m, index_t node, double out[3]) {
    const index_t n = m.n_nodes();
    const double lon = m.h_coords[node], lat = m.h_coords[n + node];
    double sl, cl, so, co;
    if (m.coord_system() == CoordinateSystem::SphericalDeg) {
        sin_cos(lat * kDeg2Rad, sl, cl);
        sin_cos(lon * kDeg2Rad, so, co);
    } else {
        sin_cos(lat, sl, cl);
        sin_cos(lon, so, co);
    }
    out[0] = cl * co;
    out[1] = cl * so;
    out[2] = sl;
}

bool maybe_rectilinear_deg(const Mesh &m) {
    if (m.coord_system() != CoordinateSystem::SphericalDeg || std::getenv("AXIS_HOST_NODE_XYZ") != nullptr) return false;
    return m.min_cell_vertices() == 4 && m.max_cell_vertices() == 4 && grid_counts(m);
}

bool node_xyz_device_eligible(const Mesh &m, const GridAxes &ax) {
    return m.coord_system() == CoordinateSystem::SphericalDeg && rectilinear_from_axes(m, ax) && std::getenv("AXIS_HOST_NODE_XYZ") == nullptr &&
           ax.xlon.size() < (std::size_t(1) << 30) && ax.xlat.size() < (std::size_t(1) << 30);
}

PinnedArray<double> axis_trig_tables(const GridAxes &ax) {
    // exact distinct coordinates where the detection kept them (else the
    // tolerance-unique ones: nodes off those values go to the host)
    const std::vector<double> &vlon = ax.xlon.empty() ? ax.ulon : ax.xlon, &vlat = ax.xlat.empty() ? ax.ulat : ax.xlat;
    const int nlon = static_cast<int>(vlon.size()), nlat = static_cast<int>(vlat.size());
    // the tables with host libm, as node_xyz_host's per-node calls would see them
    PinnedArray<double> tab;  // pinned: a pageable upload would wait on the host threads
    tab.resize(3 * static_cast<std::size_t>(nlon + nlat));
    double *ulon = tab.data(), *ulat = ulon + nlon, *slon = ulat + nlat, *clon = slon + nlon, *slat = clon + nlon, *clat = slat + nlat;
    for (int a = 0; a < nlon; ++a) {
        ulon[a] = vlon[a];
        sin_cos(vlon[a] * kDeg2Rad, slon[a], clon[a]);
    }
    for (int b = 0; b < nlat; ++b) {
        ulat[b] = vlat[b];
        sin_cos(vlat[b] * kDeg2Rad, slat[b], clat[b]);
    }
    return tab;
}

bool node_xyz_device(const Mesh &m, const GridAxes &ax, MeshGeom &g, hipStream_t s, PinnedArray<double> *tables) {
    g.dev_xyz = false;
    if (!node_xyz_device_eligible(m, ax)) return false;
    PhaseTrace tr("geometry", s);
    const index_t n = m.n_nodes();
    const int nlon = static_cast<int>((ax.xlon.empty() ? ax.ulon : ax.xlon).size()), nlat = static_cast<int>((ax.xlat.empty() ? ax.ulat : ax.xlat).size());
    PinnedArray<double> own;
    if (!tables) {
        own = axis_trig_tables(ax);
        tables = &own;
    }
    PinnedArray<double> &tab = *tables;
    DeviceBuffer<double> d_tab(tab.size());
    d_tab.upload(tab.data(), tab.size(), s);
    // a fresh buffer, moved into g.d_xyz only on success: on failure g.d_xyz
    // keeps whatever the host path uploaded into it
    DeviceBuffer<double> xyz(std::max<std::size_t>(3 * static_cast<std::size_t>(n), 3));
    DeviceBuffer<index_t> fix(kXyzFixCap);
    DeviceBuffer<int> n_fix(1);
    AXIS_HIP_CHECK(hipMemsetAsync(n_fix.get(), 0, sizeof(int), s));
    if (n > 0) {
        axis_xyz_kernel<<<grid_for(n, 256), 256, 0, s>>>(m.d_coords.get(), m.d_coords.get() + n, n, d_tab.get(), nlon, nlat, xyz.get(), fix.get(),
                                                         n_fix.get(), kXyzFixCap);
        AXIS_LAUNCH_CHECK();
    }
    int h_fix = 0;
    read_back(&h_fix, n_fix.get(), 1, s);
    if (h_fix > kXyzFixCap) return false;
    if (h_fix > 0) {
        std::vector<index_t> nodes(h_fix);
        fix.download(nodes.data(), h_fix, s);
        AXIS_HIP_CHECK(hipStreamSynchronize(s));
        std::vector<double> vals(3 * static_cast<std::size_t>(h_fix));
        for (int k = 0; k < h_fix; ++k) node_xyz_one(m, nodes[k], &vals[3 * static_cast<std::size_t>(k)]);
        DeviceBuffer<index_t> d_nodes(h_fix);
        DeviceBuffer<double> d_vals(vals.size());
        d_nodes.upload(nodes.data(), h_fix, s);
        d_vals.upload(vals.data(), vals.size(), s);
        scatter_xyz_kernel<<<grid_for(h_fix, 256), 256, 0, s>>>(d_nodes.get(), d_vals.get(), h_fix, xyz.get());
        AXIS_LAUNCH_CHECK();
        AXIS_HIP_CHECK(hipStreamSynchronize(s));
    }
    g.d_xyz = std::move(xyz);
    g.dev_xyz = true;
    tr.mark("device_xyz");
    return true;
}

// host threads for the libm angles beside a generation (AXIS_TRIG_THREADS)
static int trig_threads() {
    static const int n = [] {
        if (const char *e = std::getenv("AXIS_TRIG_THREADS"))
            if (int v = std::atoi(e); v > 0) return v;
        return std::max(1, host_threads() - 3);
    }();
    return n;
}

bool node_xyz_trig(const Mesh &m, double *d_out, hipStream_t st, const std::atomic<bool> *cancel) {
    if (m.coord_system() != CoordinateSystem::SphericalDeg || std::getenv("AXIS_HOST_NODE_XYZ") != nullptr) return false;
    const index_t n = m.n_nodes();
    if (n == 0) return true;
    PhaseTrace tr("geometry", st);
    const index_t words = (n + 63) / 64;
    // flags of lat, lon; then host libm's last bits of sin lat, cos lat, sin lon, cos lon
    DeviceBuffer<std::uint64_t> bits(6 * static_cast<std::size_t>(words));
    // (nodes on the multiples of 90 degrees: a few per thousand on a cubed sphere)
    const int full_cap = static_cast<int>(std::min<index_t>(n / 16 + kXyzFixCap, index_t(1) << 30));
    const int full_first = std::min(full_cap, 1 << 16);  // list entries fetched with the flags
    DeviceBuffer<index_t> full(full_cap);
    DeviceBuffer<int> n_full(1);
    PinnedArray<std::uint64_t> hb;
    hb.resize(6 * static_cast<std::size_t>(words) + 1);
    std::uint64_t *flags = hb.data(), *res = flags + 2 * words, *h_nfull = flags + 6 * words;
    PinnedArray<index_t> h_full;
    h_full.resize(full_cap);
    AXIS_HIP_CHECK(hipMemsetAsync(n_full.get(), 0, sizeof(int), st));
    const double *d_lon = m.d_coords.get(), *d_lat = d_lon + n;
    trig_xyz_kernel<<<grid_for(n, 256), 256, 0, st>>>(d_lon, d_lat, n, words, d_out, bits.get(), bits.get() + words, full.get(), n_full.get(),
                                                      full_cap);
    AXIS_LAUNCH_CHECK();
    // (copy kernels on this side stream, no DMA engine: kernel_copy)
    kernel_copy(flags, bits.get(), 2 * sizeof(std::uint64_t) * words, st);
    *h_nfull = 0;
    kernel_copy(h_nfull, n_full.get(), sizeof(int), st);
    kernel_copy(h_full.data(), full.get(), sizeof(index_t) * full_first, st);
    AXIS_HIP_CHECK(hipStreamSynchronize(st));
    const int nf = static_cast<int>(*h_nfull & 0xffffffffu);
    if (nf > full_cap) return false;
    if (nf > full_first) {
        kernel_copy(h_full.data() + full_first, full.get() + full_first, sizeof(index_t) * (nf - full_first), st);
        AXIS_HIP_CHECK(hipStreamSynchronize(st));
    }
    // host libm: one bit per sine and cosine of the flagged angles, every
    // value of the listed nodes (one pass: words first, then list chunks);
    // three cores stay free for the generating thread and the HIP runtime
    const double *lon = m.h_coords.data(), *lat = lon + n;
    PinnedArray<double> h_vals;
    h_vals.resize(3 * static_cast<std::size_t>(std::max(nf, 1)));
    constexpr index_t kListChunk = 64;  // list entries per word-sized work item
    const index_t list_items = (nf + kListChunk - 1) / kListChunk;
    std::atomic<bool> stopped{false};
    parallel_range(
        words + list_items,
        [&](index_t b, index_t e) {
            for (index_t w = b; w < e; ++w) {
                if (cancel && (w & 63) == 0 && cancel->load(std::memory_order_relaxed)) {
                    stopped.store(true);
                    return;
                }
                if (w >= words) {
                    const index_t k0 = (w - words) * kListChunk, k1 = std::min<index_t>(k0 + kListChunk, nf);
                    for (index_t k = k0; k < k1; ++k) node_xyz_one(m, h_full[k], &h_vals[3 * static_cast<std::size_t>(k)]);
                    continue;
                }
                std::uint64_t r[4] = {0, 0, 0, 0};
                for (int a = 0; a < 2; ++a) {
                    const double *v = a == 0 ? lat : lon;
                    for (std::uint64_t f = flags[a * words + w]; f != 0; f &= f - 1) {
                        const int t = __builtin_ctzll(f);
                        double sv, cv;
                        sin_cos(v[64 * w + t] * kDeg2Rad, sv, cv);
                        r[2 * a] |= (std::bit_cast<std::uint64_t>(sv) & 1u) << t;
                        r[2 * a + 1] |= (std::bit_cast<std::uint64_t>(cv) & 1u) << t;
                    }
                }
                for (int q = 0; q < 4; ++q) res[q * words + w] = r[q];
            }
        },
        256, trig_threads());
    if (stopped.load()) return false;
    kernel_copy(bits.get() + 2 * words, res, 4 * sizeof(std::uint64_t) * words, st);
    trig_fix_kernel<<<grid_for(n, 256), 256, 0, st>>>(d_lon, d_lat, n, words, bits.get(), bits.get() + words, bits.get() + 2 * words, d_out);
    AXIS_LAUNCH_CHECK();
    if (nf > 0) {  // the listed nodes' xyz (list order kept on the device)
        DeviceBuffer<double> d_vals(3 * static_cast<std::size_t>(nf));
        kernel_copy(d_vals.get(), h_vals.data(), sizeof(double) * 3 * nf, st);
        scatter_xyz_kernel<<<grid_for(nf, 256), 256, 0, st>>>(full.get(), d_vals.get(), nf, d_out);
        AXIS_LAUNCH_CHECK();
    }
    // (the pinned arrays go back to their cache pending the next device
    // synchronisation, so the copies above read them intact)
    tr.mark("trig_xyz");
    return true;
}

namespace {
// the current device's side stream number k (non-blocking, default priority,
// created on first use and kept until the process ends)
hipStream_t side_stream(int k) {
    static std::mutex mu;
    static std::map<std::pair<int, int>, hipStream_t> streams;
    int dev = 0;
    AXIS_HIP_CHECK(hipGetDevice(&dev));
    std::lock_guard<std::mutex> g(mu);
    auto it = streams.find({dev, k});
    if (it != streams.end()) return it->second;
    hipStream_t st = nullptr;
    AXIS_HIP_CHECK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    streams[{dev, k}] = st;
    return st;
}
}  // namespace

hipStream_t upload_stream() { return side_stream(0); }
hipStream_t rows_stream() { return side_stream(1); }

bool node_xyz_host(const Mesh &m, MeshGeom &g, const std::atomic<bool> *cancel, hipStream_t up, hipEvent_t done, double *d_out) {
    const index_t n_nodes = m.n_nodes();
    const bool deg = m.coord_system() == CoordinateSystem::SphericalDeg;
    g.h_xyz.resize(3 * static_cast<std::size_t>(n_nodes));
    g.h_xyz_deg.clear();
    if (!deg) g.h_xyz_deg.resize(3 * static_cast<std::size_t>(n_nodes));
    const double *lon = m.h_coords.data(), *lat = m.h_coords.data() + n_nodes;
    const bool upload = up != nullptr && d_out != nullptr;
    std::atomic<int> up_err{0};
    // beside a running generation (upload given), two cores stay free for the
    // generating thread and the HIP runtime's own threads
    const int max_threads = upload ? std::max(1, host_threads() - 2) : 0;
    int dev = 0;
    AXIS_HIP_CHECK(hipGetDevice(&dev));
    parallel_range(n_nodes, [&](index_t b, index_t e) {
        SinCosMemo memo;
        for (index_t i = b; i < e; ++i) {
            if (cancel && (i & 4095) == 0 && cancel->load(std::memory_order_relaxed)) return;
            geo::V3 p = deg ? memo.xyz(lon[i] * kDeg2Rad, lat[i] * kDeg2Rad) : memo.xyz(lon[i], lat[i]);
            g.h_xyz[3 * i] = p.x;
            g.h_xyz[3 * i + 1] = p.y;
            g.h_xyz[3 * i + 2] = p.z;
            if (!deg) {
                geo::V3 q = memo.xyz(lon[i] * kDeg2Rad, lat[i] * kDeg2Rad);
                g.h_xyz_deg[3 * i] = q.x;
                g.h_xyz_deg[3 * i + 1] = q.y;
                g.h_xyz_deg[3 * i + 2] = q.z;
            }
        }
        // this slice to the device now (errors are reported after the join:
        // a worker thread must not throw)
        if (upload && e > b) {  // (a copy kernel, as the side stream's other copies)
            try {
                AXIS_HIP_CHECK(hipSetDevice(dev));  // (a pool thread: the generating thread's device)
                kernel_copy(d_out + 3 * b, g.h_xyz.data() + 3 * b, 3 * sizeof(double) * static_cast<std::size_t>(e - b), up);
            } catch (...) {
                up_err.store(1);
            }
        }
    }, 4096, max_threads);
    const bool complete = !(cancel && cancel->load());
    g.xyz_ready = complete;
    if (up_err.load() != 0) throw RuntimeError("node xyz: a slice upload failed");
    if (!upload || !complete) return false;
    AXIS_HIP_CHECK(hipEventRecord(done, up));
    return true;
}

void prepare_geometry(const Mesh &m, MeshGeom &g, hipStream_t s) {
    const index_t n_nodes = m.n_nodes(), n_cells = m.n_cells();
    g.n_nodes = n_nodes;
    g.n_cells = n_cells;
    const bool deg = m.coord_system() == CoordinateSystem::SphericalDeg;
    PhaseTrace tr("geometry", s);
    // node xyz with host libm (extract_cell_polygon_spherical, weight_generator.cpp:462-470)
    // (or built on the device, node_xyz_device)
    if (!g.dev_xyz && !g.d_xyz_ready && !g.xyz_ready && deg) {
        DeviceBuffer<double> xyz(std::max<std::size_t>(3 * static_cast<std::size_t>(n_nodes), 3));
        if (node_xyz_trig(m, xyz.get(), s)) {
            g.d_xyz = std::move(xyz);
            g.dev_xyz = true;
        }
    }
    if (!g.dev_xyz && !g.d_xyz_ready) {
        if (!g.xyz_ready) node_xyz_host(m, g);
        g.d_xyz.resize(std::max<std::size_t>(g.h_xyz.size(), 3));
        g.d_xyz.upload(g.h_xyz.data(), g.h_xyz.size(), s);
    }
    const PinnedArray<double> &h_xyz_deg = g.h_xyz_deg;
    tr.mark("host_sincos");
    DeviceBuffer<double> d_xyz_deg;
    if (!deg) {
        d_xyz_deg.resize(std::max<std::size_t>(h_xyz_deg.size(), 3));
        d_xyz_deg.upload(h_xyz_deg.data(), h_xyz_deg.size(), s);
    }
    // trig cache of a regular grid (trig_cache.hpp:248-261)
    if (g.regular) {
        const index_t nl = g.ni + 1, na = g.nj + 1;
        g.h_trig.resize(2 * (nl + na));
        for (index_t i = 0; i < nl; ++i) {
            double r = (g.lon_min + static_cast<double>(i) * g.delta_lon) * kDeg2Rad;
            sin_cos(r, g.h_trig[i], g.h_trig[nl + i]);
        }
        for (index_t j = 0; j < na; ++j) {
            double r = (g.lat_min + static_cast<double>(j) * g.delta_lat) * kDeg2Rad;
            sin_cos(r, g.h_trig[2 * nl + j], g.h_trig[2 * nl + na + j]);
        }
        g.d_trig.resize(g.h_trig.size());
        g.d_trig.upload(g.h_trig.data(), g.h_trig.size(), s);
    }
    const index_t nc = std::max<index_t>(n_cells, 1);
    g.d_area.resize(nc);
    g.d_flags.resize(nc);
    g.d_cap.resize(4 * nc);
    g.d_box.resize(6 * nc);
    DeviceBuffer<index_t> recheck(nc);
    DeviceBuffer<int> n_recheck(1);
    AXIS_HIP_CHECK(hipMemsetAsync(n_recheck.get(), 0, sizeof(int), s));
    if (n_cells > 0) {
        cell_prep_kernel<<<grid_for(n_cells, kCellBlock), kCellBlock, 0, s>>>(
            n_cells, m.d_offsets.get(), m.d_indices.get(), g.d_xyz.get(), deg ? g.d_xyz.get() : d_xyz_deg.get(),
            m.has_areas() ? m.d_areas.get() : nullptr, m.has_mask() ? m.d_mask.get() : nullptr, g.d_area.get(), g.d_flags.get(), g.d_cap.get(),
            g.d_box.get(), recheck.get(), n_recheck.get());
        AXIS_LAUNCH_CHECK();
    }
    tr.mark("upload_trig_cellprep");
    int h_n = 0;
    read_back(&h_n, n_recheck.get(), 1, s);
    if (h_n > 0) {
        // replay near-zero-area cells with host libm (the reference's bits)
        std::vector<index_t> cells(h_n);
        recheck.download(cells.data(), h_n, s);
        std::vector<double> area(h_n);
        std::vector<std::int32_t> flags(h_n);
        DeviceBuffer<double> tmp_area(h_n);
        DeviceBuffer<std::int32_t> tmp_flags(h_n);
        AXIS_HIP_CHECK(hipStreamSynchronize(s));
        const double *xyz_deg = deg ? g.h_xyz.data() : h_xyz_deg.data();
        std::vector<double> own_xyz;  // device-built xyz: the cell's nodes with host libm
        std::vector<index_t> own_idx;
        for (int k = 0; k < h_n; ++k) {
            index_t c = cells[k];
            index_t b = m.h_offsets[c];
            int nv = static_cast<int>(std::min<index_t>(m.h_offsets[c + 1] - b, 32));
            double fan_deg;
            int cls;
            if (g.dev_xyz) {  // (degree meshes only)
                own_xyz.resize(3 * static_cast<std::size_t>(nv));
                own_idx.resize(nv);
                for (int v = 0; v < nv; ++v) {
                    node_xyz_one(m, m.h_indices[b + v], &own_xyz[3 * static_cast<std::size_t>(v)]);
                    own_idx[v] = v;
                }
                cls = host_classify(own_xyz.data(), own_idx.data(), nv, fan_deg);
            } else {
                cls = host_classify(xyz_deg, m.h_indices.data() + b, nv, fan_deg);
            }
            double a;
            if (m.has_areas())
                a = m.h_areas[c];
            else if (deg)
                a = fan_deg;
            else {
                double dummy;
                host_classify(g.h_xyz.data(), m.h_indices.data() + b, nv, dummy);
                a = dummy;
            }
            area[k] = a;
            flags[k] = (cls != 0 ? 1 : 0) | ((m.has_mask() && m.h_mask[c] == 0) ? 2 : 0) | (a <= 0.0 ? 4 : 0);
        }
        // the replayed values in one upload each, scattered on the device
        DeviceBuffer<index_t> d_cells(h_n);
        d_cells.upload(cells.data(), h_n, s);
        tmp_area.upload(area.data(), h_n, s);
        tmp_flags.upload(flags.data(), h_n, s);
        scatter_recheck_kernel<<<grid_for(h_n, 256), 256, 0, s>>>(d_cells.get(), tmp_area.get(), tmp_flags.get(), h_n, g.d_area.get(), g.d_flags.get());
        AXIS_LAUNCH_CHECK();
        AXIS_HIP_CHECK(hipStreamSynchronize(s));
    }
}

namespace {
__global__ void cell_key_kernel(index_t n_cells, const index_t *off, const index_t *ind, const double *coords, index_t n_nodes, double to_rad,
                                std::uint32_t *key) {
    const index_t c = static_cast<index_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (c >= n_cells) return;
    double x = 0.0, y = 0.0, z = 0.0;
    for (index_t q = off[c]; q < off[c + 1]; ++q) {
        const index_t v = ind[q];
        double so, co, sa, ca;
        sincos(coords[v] * to_rad, &so, &co);
        sincos(coords[n_nodes + v] * to_rad, &sa, &ca);
        x += ca * co;
        y += ca * so;
        z += sa;
    }
    const double len = sqrt(x * x + y * y + z * z);
    const double inv = len > 0.0 ? 1.0 / len : 0.0;
    key[c] = morton30(static_cast<float>(x * inv), static_cast<float>(y * inv), static_cast<float>(z * inv));
}
}  // namespace

void cell_morton_keys(const Mesh &m, DeviceBuffer<std::uint32_t> &key, hipStream_t s) {
    key.resize(std::max<index_t>(m.n_cells(), 1));
    if (m.n_cells() == 0) return;
    const double to_rad = m.coord_system() == CoordinateSystem::SphericalDeg ? kDeg2Rad : 1.0;
    cell_key_kernel<<<grid_for(m.n_cells(), 256), 256, 0, s>>>(m.n_cells(), m.d_offsets.get(), m.d_indices.get(), m.d_coords.get(), m.n_nodes(),
                                                              to_rad, key.get());
    AXIS_LAUNCH_CHECK();
}

}  // namespace detail
}  // namespace axis_hip
