// conservative.hip -- first-order conservative weights on great circles.
//
// Restates the reference host path generate_conservative /
// generate_conservative_impl<3> (libs/axis/src/solver/weight_generator.cpp:
// 3246-3302, 3304-3770) as a GPU pipeline:
//
//   detection ........ unique_axes / regular_from_axes        (mesh.hip)
//   per-cell pass .... areas, degenerate flags, caps, boxes   (mesh.hip)
//   broad phase ...... float box overlaps: through the source's cell grid
//                      when it is a rectilinear degree grid (gridquery.hip),
//                      else an LBVH (lbvh.hip); the same candidate sets
//   overlap pass ..... cap filter per (dst, candidate) pair, compacting the
//                      survivors; then one lane per survivor: great-circle
//                      Sutherland-Hodgman clip with the subject polygon in
//                      lane-private LDS columns, FP64 fan area
//   arbitration ...... the few pairs whose outcome hinges on libm ulps
//                      (cap-test ties, atan2 cancellation, >CAP vertices)
//                      are recomputed on the host with the reference's libm
//   CSR assembly ..... per-row nonzero counts, row_ptr by a scan, each row's
//                      nonzeros sorted by column where they lie (candidates
//                      are grouped by row), weights w = overlap / area_dst,
//                      frac_b (column order), FracArea scaling, frac_a
//                      (row order, by a stable column sort), clamps
//
// The sparsity pattern equals the reference's bit for bit: an entry exists
// iff the clipped area is > 0, and that decision depends only on + - * /
// sqrt (identical on both sides under -ffp-contract=off) except where the
// arbitration hands it to host libm.
#include <hip/hip_runtime.h>

#include <atomic>
#include <future>

#include <algorithm>
#include <chrono>
#include <cstring>
#include <cmath>
#include <limits>
#include <map>
#include <string>

#include "axis_hip.hpp"
#include "geometry.hpp"
#include "host_util.hpp"
#include "proposer.hpp"

namespace axis_hip::detail {

using geo::PolyRef;
using geo::V3;

namespace {

constexpr int kWave = 64;

// Relative band around the cap-test threshold inside which the GPU's acos
// could disagree with host libm; such pairs are clipped and re-decided.
constexpr double kCapTieRel = 1.0e-12;
// Fan areas whose terms cancel below this fraction of their magnitude sum
// are recomputed with host atan2.
constexpr double kCancelRel = 1.0e-3;

struct Side {
    const index_t *off;
    const index_t *ind;
    const double *xyz;
    const double *trig;  // non-null for a regular lat-lon mesh
    index_t ni, nlon, nlat;
    const std::int32_t *flags;
    const double *cap;
    const double *area;
};

__device__ inline V3 trig_node(const Side &m, index_t a, index_t b) {
    const double *slon = m.trig, *clon = m.trig + m.nlon, *slat = m.trig + 2 * m.nlon, *clat = m.trig + 2 * m.nlon + m.nlat;
    double cl = clat[b];
    return V3{cl * clon[a], cl * slon[a], slat[b]};
}

// vertex k of cell c: trig cache corners BL, BR, TR, TL
// (weight_generator.cpp:483-520) or mesh nodes
__device__ inline V3 cell_vertex(const Side &m, index_t c, int k) {
    if (m.trig) {
        // (32-bit division: cell counts stay below 2^31, and a 64-bit one
        // costs many times more instructions)
        const std::uint32_t cu = static_cast<std::uint32_t>(c), ni = static_cast<std::uint32_t>(m.ni);
        const std::uint32_t cj = cu / ni, ci = cu - cj * ni;
        const index_t a = (k == 1 || k == 2) ? ci + 1 : ci;
        const index_t b = (k >= 2) ? cj + 1 : cj;
        return trig_node(m, a, b);
    }
    index_t nd = m.ind[m.off[c] + k];
    return V3{m.xyz[3 * nd], m.xyz[3 * nd + 1], m.xyz[3 * nd + 2]};
}

__device__ inline int cell_nverts(const Side &m, index_t c) {
    if (m.trig) return 4;
    index_t nv = m.off[c + 1] - m.off[c];
    return static_cast<int>(nv < 32 ? nv : 32);
}

// A pair for host libm: its index in arb (the count in n_arb), and (pair,
// destination row, source cell) in trip while the count is below kTripCap,
// so the host fetches what it needs with the count in one readback.
constexpr int kTripCap = 1 << 16;

__device__ inline void queue_for_host(index_t p, index_t j, index_t i, std::int32_t *arb, int *n_arb, std::int32_t *trip) {
    const int k = atomicAdd(n_arb, 1);
    arb[k] = static_cast<std::int32_t>(p);
    if (k < kTripCap) {
        trip[3 * k] = static_cast<std::int32_t>(p);
        trip[3 * k + 1] = static_cast<std::int32_t>(j);
        trip[3 * k + 2] = static_cast<std::int32_t>(i);
    }
}

template <int CAP>
__global__ void __launch_bounds__(kWave) overlap_kernel(index_t p0, index_t p1, const std::int32_t *pair_row, const std::int32_t *pair_src,
                                                        Side src, Side dst, double *overlap, std::int32_t *arb, int *n_arb, int *n_tie,
                                                        std::int32_t *trip) {
    __shared__ double buf[3][CAP][kWave];  // one lane-private polygon per lane, clipped in place
    const int lane = threadIdx.x;
    const index_t p = p0 + static_cast<index_t>(blockIdx.x) * kWave + lane;
    if (p >= p1) return;
    const index_t j = pair_row[p];
    const index_t i = pair_src[p];
    double ov = 0.0;
    bool to_host = false;
    // masked / degenerate / area <= 0 (weight_generator.cpp:3618-3634, 3650-3660)
    if (((dst.flags[j] | src.flags[i]) & 7) == 0) {
        // spherical cap early reject (spherical_cap_filter.hpp:190-197)
        const double *cs = src.cap + 4 * i, *cd = dst.cap + 4 * j;
        double c = cs[0] * cd[0] + cs[1] * cd[1] + cs[2] * cd[2];
        c = fmin(1.0, fmax(-1.0, c));
        const double ang = acos(c);
        const double sum_r = cs[3] + cd[3];
        const double tie = kCapTieRel * sum_r + 1.0e-300;
        if (ang <= sum_r + tie) {
            if (ang > sum_r - tie) {
                to_host = true;
                atomicAdd(n_tie, 1);
            }
            PolyRef A{&buf[0][0][lane], &buf[1][0][lane], &buf[2][0][lane], kWave};
            const int ns = cell_nverts(src, i);
            bool overflow = ns > CAP;
            if (!overflow) {
                for (int k = 0; k < ns; ++k) A.set(k, cell_vertex(src, i, k));
                const int nc = cell_nverts(dst, j);
                auto clipv = [&](int k) { return cell_vertex(dst, j, k); };
                int n = geo::clip_polygon_slots(A, ns, clipv, nc, CAP, &overflow);
                if (!overflow && n >= 3) {
                    double abs_sum;
                    bool zero_num;
                    ov = geo::polygon_area(A, n, &abs_sum, &zero_num);
                    if (!zero_num && ov < kCancelRel * abs_sum) to_host = true;
                }
            }
            if (overflow) to_host = true;
        }
    }
    overlap[p] = ov;
    if (to_host) queue_for_host(p, j, i, arb, n_arb, trip);
}

// Two-kernel form of overlap_kernel for pair ranges below 2^31: the flags and
// cap tests first (about 40% of the LBVH's box candidates fail the cap test),
// compacting the surviving pairs, then the clipping over the survivors only,
// so that no clipping wave carries lanes that have nothing to do.  Each
// pair's arithmetic is that of overlap_kernel.  The compaction is per chunk
// of kSelChunk pairs (survivors at the front of the chunk's own range of
// `live`, in pair order, their count in `n_live`): no global atomics.
constexpr std::uint32_t kTieBit = 0x80000000u;
constexpr int kSelThreads = 256;
constexpr int kSelIter = 16;
constexpr int kSelChunk = kSelThreads * kSelIter;  // pairs per chunk
constexpr int kSelWaves = kSelThreads / kWave;

__global__ void __launch_bounds__(kSelThreads) cap_select_kernel(index_t p0, index_t p1, const std::int32_t *pair_row, const std::int32_t *pair_src,
                                                                 Side src, Side dst, double *overlap, std::uint32_t *live, std::int32_t *n_live,
                                                                 int *n_tie) {
    __shared__ unsigned wc[2][kSelWaves];
    const int tid = threadIdx.x, lane = tid & (kWave - 1), wave = tid / kWave;
    const index_t cb = p0 + static_cast<index_t>(blockIdx.x) * kSelChunk;
    std::uint32_t *out = live + static_cast<index_t>(blockIdx.x) * kSelChunk;
    unsigned run = 0;
    for (int it = 0; it < kSelIter; ++it) {
        const index_t p = cb + it * kSelThreads + tid;
        bool keep = false, tie = false;
        if (p < p1) {
            const index_t j = pair_row[p];
            const index_t i = pair_src[p];
            if (((dst.flags[j] | src.flags[i]) & 7) == 0) {
                const double *cs = src.cap + 4 * i, *cd = dst.cap + 4 * j;
                double c = cs[0] * cd[0] + cs[1] * cd[1] + cs[2] * cd[2];
                c = fmin(1.0, fmax(-1.0, c));
                const double ang = acos(c);
                const double sum_r = cs[3] + cd[3];
                const double tw = kCapTieRel * sum_r + 1.0e-300;
                if (ang <= sum_r + tw) {
                    keep = true;
                    if (ang > sum_r - tw) {
                        tie = true;
                        atomicAdd(n_tie, 1);
                    }
                }
            }
            if (!keep) overlap[p] = 0.0;
        }
        const unsigned long long m = __ballot(keep);
        if (lane == 0) wc[it & 1][wave] = static_cast<unsigned>(__popcll(m));
        __syncthreads();  // (the double buffer lets one barrier per step suffice)
        unsigned off = run, tot = 0;
#pragma unroll
        for (int w = 0; w < kSelWaves; ++w) {
            const unsigned cw = wc[it & 1][w];
            if (w < wave) off += cw;
            tot += cw;
        }
        if (keep) out[off + static_cast<unsigned>(__popcll(m & ((1ull << lane) - 1ull)))] = static_cast<std::uint32_t>(p - p0) | (tie ? kTieBit : 0u);
        run += tot;
    }
    if (tid == 0) n_live[blockIdx.x] = static_cast<std::int32_t>(run);
}

template <int CAP>
__global__ void __launch_bounds__(kWave) clip_live_kernel(index_t p0, const std::uint32_t *live, const std::int32_t *n_live, unsigned n_chunks,
                                                          const std::int32_t *pair_row, const std::int32_t *pair_src, Side src, Side dst,
                                                          double *overlap, std::int32_t *arb, int *n_arb, std::int32_t *trip) {
    __shared__ double buf[3][CAP][kWave];
    const int lane = threadIdx.x;
    // block b clips wave b / n_chunks of chunk b % n_chunks: the blocks with
    // work come first, and consecutive blocks take distant parts of the pair
    // list (at 0.25deg -> C384, 1.06 ms against 1.65 ms for chunk-major
    // order and 1.6 ms for the one-kernel form)
    const unsigned chunk = blockIdx.x % n_chunks;
    const int q = static_cast<int>((blockIdx.x / n_chunks) * kWave) + lane;
    if (q >= n_live[chunk]) return;
    const std::uint32_t e = live[static_cast<index_t>(chunk) * kSelChunk + q];
    const index_t p = p0 + static_cast<index_t>(e & ~kTieBit);
    bool to_host = (e & kTieBit) != 0;
    const index_t j = pair_row[p];
    const index_t i = pair_src[p];
    double ov = 0.0;
    PolyRef A{&buf[0][0][lane], &buf[1][0][lane], &buf[2][0][lane], kWave};
    const int ns = cell_nverts(src, i);
    bool overflow = ns > CAP;
    if (!overflow) {
        const int nc = cell_nverts(dst, j);
        int n;
        if (ns == 4 && nc == 4) {
            // two quads: all eight vertices loaded together, once (the clip
            // cell's vertices then come from registers)
            const V3 s0 = cell_vertex(src, i, 0), s1 = cell_vertex(src, i, 1), s2 = cell_vertex(src, i, 2), s3 = cell_vertex(src, i, 3);
            const V3 q0 = cell_vertex(dst, j, 0), q1 = cell_vertex(dst, j, 1), q2 = cell_vertex(dst, j, 2), q3 = cell_vertex(dst, j, 3);
            A.set(0, s0);
            A.set(1, s1);
            A.set(2, s2);
            A.set(3, s3);
            auto clipq = [&](int k) { return k == 0 ? q0 : (k == 1 ? q1 : (k == 2 ? q2 : q3)); };
            n = geo::clip_polygon_slots(A, 4, clipq, 4, CAP, &overflow);
        } else {
            for (int k = 0; k < ns; ++k) A.set(k, cell_vertex(src, i, k));
            auto clipv = [&](int k) { return cell_vertex(dst, j, k); };
            n = geo::clip_polygon_slots(A, ns, clipv, nc, CAP, &overflow);
        }
        if (!overflow && n >= 3) {
            double abs_sum;
            bool zero_num;
            ov = geo::polygon_area(A, n, &abs_sum, &zero_num);
            if (!zero_num && ov < kCancelRel * abs_sum) to_host = true;
        }
    }
    if (overflow) to_host = true;
    overlap[p] = ov;
    if (to_host) queue_for_host(p, j, i, arb, n_arb, trip);
}

// Rows the proposer's round must hand to the LBVH: rejected by the guard, or
// accepted but with overlaps (candidate order) not summing to the cell area
// within `tol` relative (1e-12 unless the caller sets it).  Cells excluded from weights never need a redo.
__global__ void redo_rows_kernel(const index_t *cand_off, index_t n_dst, const double *overlap, const double *area, const std::int32_t *flags,
                                 const std::int32_t *accepted, double tol, index_t *redo, unsigned long long *n_guard, unsigned long long *n_area) {
    index_t j = static_cast<index_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (j >= n_dst) return;
    index_t r = 0;
    if ((flags[j] & 7) == 0) {
        if (!accepted[j]) {
            r = 1;
            atomicAdd(n_guard, 1ull);
        } else {
            double sum = 0.0;
            for (index_t e = cand_off[j]; e < cand_off[j + 1]; ++e) sum += overlap[e];
            if (!(fabs(sum - area[j]) <= tol * area[j])) {
                r = 1;
                atomicAdd(n_area, 1ull);
            }
        }
    }
    redo[j] = r;
}

__global__ void redo_list_kernel(const index_t *redo, const index_t *pos, index_t n_dst, const float *box, std::int32_t *rows, float *rbox) {
    index_t j = static_cast<index_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (j >= n_dst || !redo[j]) return;
    const index_t q = pos[j];
    rows[q] = static_cast<std::int32_t>(j);
    for (int a = 0; a < 6; ++a) rbox[6 * q + a] = box[6 * j + a];
}

// round-1 pairs of redone rows drop out (overlap 0 -> no entry)
__global__ void zero_rows_kernel(const std::int32_t *rows, index_t n_rows, const index_t *cand_off, double *overlap) {
    index_t q = static_cast<index_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (q >= n_rows) return;
    const index_t j = rows[q];
    for (index_t e = cand_off[j]; e < cand_off[j + 1]; ++e) overlap[e] = 0.0;
}

// round-2 pairs: dst row of each LBVH candidate of redone row q
__global__ void fill_rows_mapped_kernel(const index_t *off, index_t n_rows, const std::int32_t *rows, std::int32_t *pair_row) {
    index_t q = static_cast<index_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (q >= n_rows) return;
    for (index_t e = off[q]; e < off[q + 1]; ++e) pair_row[e] = rows[q];
}

// dst row of every candidate pair: a block stages the offsets of its 256 rows
// and writes their pair range contiguously, each pair's row found by binary
// search in LDS
__global__ void __launch_bounds__(256) fill_rows_kernel(const index_t *cand_off, index_t n_dst, std::int32_t *pair_row) {
    __shared__ index_t so[257];
    const index_t r0 = static_cast<index_t>(blockIdx.x) * 256;
    const int nr = static_cast<int>(min<index_t>(256, n_dst - r0));
    for (int k = threadIdx.x; k <= nr; k += 256) so[k] = cand_off[r0 + k];
    __syncthreads();
    const index_t b = so[0], e = so[nr];
    for (index_t p = b + threadIdx.x; p < e; p += 256) {
        int lo = 0, hi = nr;  // so[lo] <= p < so[hi]
        while (hi - lo > 1) {
            const int mid = (lo + hi) >> 1;
            if (so[mid] <= p)
                lo = mid;
            else
                hi = mid;
        }
        pair_row[p] = static_cast<std::int32_t>(r0 + lo);
    }
}

// (pair, dst row, src cell) of the pairs handed to the host, in one buffer
__global__ void gather_arbitrated_kernel(const std::int32_t *arb, int n, const std::int32_t *pair_row, const std::int32_t *cand_src,
                                         std::int32_t *out) {
    int k = static_cast<int>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const std::int32_t p = arb[k];
    out[3 * k] = p;
    out[3 * k + 1] = pair_row[p];
    out[3 * k + 2] = cand_src[p];
}

__global__ void scatter_overlaps_kernel(const std::int32_t *idx, const double *vals, int n, double *overlap) {
    int k = static_cast<int>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (k < n) overlap[idx[k]] = vals[k];
}

__global__ void nz_flag_kernel(const double *overlap, index_t n, index_t *flag) {
    index_t p = static_cast<index_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (p < n) flag[p] = overlap[p] > 0.0 ? 1 : 0;
}

// compact nonzero pairs, keyed (major << minor_bits | minor)
__global__ void compact_kernel(const double *overlap, const index_t *pos, index_t n, const std::int32_t *pair_row, const std::int32_t *pair_src,
                               int col_bits, std::uint64_t *key, std::uint32_t *perm, double *ov_out) {
    index_t p = static_cast<index_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (p >= n || !(overlap[p] > 0.0)) return;
    index_t q = pos[p];
    key[q] = (static_cast<std::uint64_t>(pair_row[p]) << col_bits) | static_cast<std::uint64_t>(static_cast<std::uint32_t>(pair_src[p]));
    perm[q] = static_cast<std::uint32_t>(q);
    ov_out[q] = overlap[p];
}

// sorted (row, col) keys -> col, overlap, row; row_ptr from the row boundaries
// (rows without entries get the offset of the next entry)
__global__ void unpack_sorted_kernel(const std::uint64_t *key, const std::uint32_t *perm, const double *ov_in, index_t nnz, int col_bits,
                                     index_t n_dst, index_t *col, double *ov, index_t *row_ptr) {
    index_t e = static_cast<index_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (e >= nnz) return;
    std::uint64_t k = key[e];
    index_t r = static_cast<index_t>(k >> col_bits);
    col[e] = static_cast<index_t>(k & ((std::uint64_t(1) << col_bits) - 1));
    ov[e] = ov_in[perm[e]];
    const index_t rp = e == 0 ? -1 : static_cast<index_t>(key[e - 1] >> col_bits);
    for (index_t q = rp + 1; q <= r; ++q) row_ptr[q] = e;
    if (e == nnz - 1)
        for (index_t q = r + 1; q <= n_dst; ++q) row_ptr[q] = nnz;
}

// per row: w = ov / area_dst (>= 0), frac_b = Σ ov / area_dst in column order,
// FracArea division by the unclamped frac_b, clamp (weight_generator.cpp:3682-3724)
__global__ void __launch_bounds__(256) row_weights_kernel(const index_t *row_ptr, index_t n_dst, const double *area_dst, const double *ov, double *w,
                                                          double *frac_b, int frac_area, int *max_row) {
    int len = 0;
    for (index_t r = static_cast<index_t>(blockIdx.x) * 256 + threadIdx.x; r < n_dst; r += static_cast<index_t>(gridDim.x) * 256) {
        const index_t b = row_ptr[r], e = row_ptr[r + 1];
        const double a = area_dst[r];
        double fb = 0.0;
        for (index_t k = b; k < e; ++k) {
            double x = ov[k] / a;
            w[k] = fmax(x, 0.0);
            fb += x;
        }
        if (frac_area && fb > 0.0)
            for (index_t k = b; k < e; ++k) w[k] /= fb;
        frac_b[r] = fmin(fb, 1.0);
        len = max(len, static_cast<int>(e - b));
    }
    // longest row: wave, then workgroup reduction, one atomic per workgroup
    for (int d = 32; d >= 1; d >>= 1) len = max(len, __shfl_xor(len, d, 64));
    __shared__ int part[4];
    if ((threadIdx.x & 63) == 0) part[threadIdx.x / 64] = len;
    __syncthreads();
    if (threadIdx.x == 0) {
        len = max(max(part[0], part[1]), max(part[2], part[3]));
        if (len > 0) atomicMax(max_row, len);
    }
}

__global__ void column_keys_kernel(const index_t *col, index_t nnz, std::uint32_t *key, std::uint32_t *perm) {
    index_t e = static_cast<index_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (e >= nnz) return;
    key[e] = static_cast<std::uint32_t>(col[e]);
    perm[e] = static_cast<std::uint32_t>(e);
}

// frac_a: Σ ov / area_src over each source column's entries in row order.
// The CSR entries are in row order and the column sort is stable, so each
// column's run lists its rows in order; a thread owns the run that starts at
// its position.
__global__ void frac_a_kernel(const std::uint32_t *key, const std::uint32_t *perm, index_t nnz, const double *ov, const double *area_src,
                              double *frac_a) {
    index_t e = static_cast<index_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (e >= nnz) return;
    const std::uint32_t c = key[e];
    if (e > 0 && key[e - 1] == c) return;
    const double a = area_src[c];
    double s = 0.0;
    for (index_t k = e; k < nnz && key[k] == c; ++k) s += ov[perm[k]] / a;
    frac_a[c] = fmin(s, 1.0);
}

// frac_a without a global sort.  Each source column's entries are gathered
// by counting (column counts whose atomics give each entry a rank within its
// column, their scan, a scatter to offset + rank that leaves each column's
// entry indices in arbitrary order), then each column sorts its few
// indices and sums in ascending index, which is the row order of
// frac_a_kernel's stable sort: the same additions in the same order.  The
// entries are the CSR's, or, where the candidate pairs are grouped by row
// with rows ascending, the candidates of positive overlap (a row holds a
// column at most once, so ascending candidate index is ascending row too).
constexpr int kColNet = 16;       // columns of up to kColNet entries: one thread, a sorting network
constexpr int kColLds = 4096;     // longer columns: one workgroup, a bitonic sort in LDS
constexpr int kColBlock = 256;

// keep (optional): entries of keep[e] > 0 only.  Each entry takes the
// column's count before its own increment as its rank (an arbitrary order
// within the column: the column's entries are sorted afterwards).
template <class C>
__global__ void col_count_kernel(const C *col, const double *keep, index_t n, unsigned long long *cnt, std::uint32_t *rank) {
    const index_t e = static_cast<index_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (e < n && (!keep || keep[e] > 0.0)) rank[e] = static_cast<std::uint32_t>(atomicAdd(&cnt[col[e]], 1ull));
}

template <class C>
__global__ void col_scatter_kernel(const C *col, const double *keep, index_t n, const index_t *cp, const std::uint32_t *rank, std::uint32_t *cs) {
    const index_t e = static_cast<index_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (e >= n || (keep && !(keep[e] > 0.0))) return;
    cs[cp[col[e]] + rank[e]] = static_cast<std::uint32_t>(e);
}

__global__ void __launch_bounds__(kColBlock) frac_a_short_kernel(const index_t *cp, const std::uint32_t *cs, index_t n_src, const double *ov,
                                                                 const double *area_src, double *frac_a, std::int32_t *long_cols, int *n_long) {
    const index_t c = static_cast<index_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (c >= n_src) return;
    const index_t b = cp[c];
    const index_t n = cp[c + 1] - b;
    if (n > kColNet) {
        long_cols[atomicAdd(n_long, 1)] = static_cast<std::int32_t>(c);
        return;
    }
    if (n == 0) {
        frac_a[c] = 0.0;
        return;
    }
    std::uint32_t v[kColNet];
#pragma unroll
    for (int k = 0; k < kColNet; ++k) v[k] = k < n ? cs[b + k] : 0xFFFFFFFFu;
    // bitonic sorting network, ascending (constant indices: the values stay in registers)
#pragma unroll
    for (int k = 2; k <= kColNet; k <<= 1)
#pragma unroll
        for (int j = k >> 1; j > 0; j >>= 1)
#pragma unroll
            for (int i = 0; i < kColNet; ++i) {
                const int l = i ^ j;
                if (l > i) {
                    const std::uint32_t x = v[i], y = v[l];
                    const bool up = (i & k) == 0;
                    v[i] = up ? min(x, y) : max(x, y);
                    v[l] = up ? max(x, y) : min(x, y);
                }
            }
    const double a = area_src[c];
    double s = 0.0;
#pragma unroll
    for (int k = 0; k < kColNet; ++k)
        if (k < n) s += ov[v[k]] / a;
    frac_a[c] = fmin(s, 1.0);
}

// one workgroup per listed column; a column beyond kColLds entries sets
// *overflow (its frac_a is then taken from the global sort)
__global__ void __launch_bounds__(kColBlock) frac_a_long_kernel(const index_t *cp, const std::uint32_t *cs, const std::int32_t *long_cols,
                                                                const int *n_long, const double *ov, const double *area_src, double *frac_a,
                                                                int *overflow) {
    __shared__ std::uint32_t sv[kColLds];
    const int nl = *n_long;
    for (int li = blockIdx.x; li < nl; li += gridDim.x) {
        const index_t c = long_cols[li];
        const index_t b = cp[c];
        const int n = static_cast<int>(cp[c + 1] - b);
        if (n > kColLds) {
            if (threadIdx.x == 0) *overflow = 1;
            continue;
        }
        int P = 1;
        while (P < n) P <<= 1;
        for (int k = threadIdx.x; k < P; k += kColBlock) sv[k] = k < n ? cs[b + k] : 0xFFFFFFFFu;
        __syncthreads();
        for (int k = 2; k <= P; k <<= 1)
            for (int j = k >> 1; j > 0; j >>= 1) {
                for (int i = threadIdx.x; i < P; i += kColBlock) {
                    const int l = i ^ j;
                    if (l > i) {
                        const std::uint32_t x = sv[i], y = sv[l];
                        const bool up = (i & k) == 0;
                        if (up ? x > y : x < y) {
                            sv[i] = y;
                            sv[l] = x;
                        }
                    }
                }
                __syncthreads();
            }
        if (threadIdx.x == 0) {
            const double a = area_src[c];
            double s = 0.0;
            for (int k = 0; k < n; ++k) s += ov[sv[k]] / a;
            frac_a[c] = fmin(s, 1.0);
        }
        __syncthreads();  // (sv is reused by the next column)
    }
}

// first unmapped destination under UnmappedAction::Error
__global__ void unmapped_kernel(const index_t *row_ptr, index_t n_dst, const std::int32_t *dflags, unsigned long long *first) {
    index_t r = static_cast<index_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (r >= n_dst) return;
    int f = dflags[r];
    if (f & 3) return;  // masked or degenerate: skipped by the reference
    if ((f & 4) || row_ptr[r + 1] == row_ptr[r]) atomicMin(first, static_cast<unsigned long long>(r));
}

// ---- CSR assembly for pairs grouped by destination row (every path but a
// proposer round 2 that appended pairs): after compaction the nonzero overlaps
// of a row are contiguous, so the (row, col) order is a per-row column sort.
// Weights, frac_b and the longest row come out of the same pass
// (weight_generator.cpp:3682-3724: w = max(ov / area_b, 0), frac_b = Σ ov /
// area_b in column order, FracArea division by the unclamped sum).
constexpr int kShortRow = 16;     // sorted by one thread
constexpr int kShortBatch = 4;    // candidates a thread loads at once
constexpr int kLongRowMax = 4096; // sorted by one workgroup in LDS
constexpr int kLongRowBlock = 256;
constexpr int kHugeRowMax = 8192; // sorted by one 1024-thread workgroup in 96 KiB of LDS
constexpr int kHugeRowBlock = 1024;
constexpr int kMidRow = 64;        // sorted by one wave in a small LDS tile

constexpr int kHistMax = 16;  // row-length histogram: lengths 0 .. kHistMax, then longer

// Nonzero overlaps per destination row (row_ptr after a scan): one thread per
// pair.  A row's pairs are contiguous and rows ascend, so each wave splits
// into runs of one row; the first lane of a run adds the run's count.
__global__ void __launch_bounds__(256) nz_count_kernel(const double *overlap, const std::int32_t *pair_row, index_t n, unsigned long long *rnz) {
    const index_t p = static_cast<index_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    const int lane = static_cast<int>(threadIdx.x & 63);
    const bool in = p < n;
    const std::int32_t r = in ? pair_row[p] : -1;
    const bool nz = in && overlap[p] > 0.0;
    const std::int32_t prev = __shfl_up(r, 1);
    const bool head = in && (lane == 0 || prev != r);
    const std::uint64_t heads = __ballot(head), nzm = __ballot(nz);
    if (!head) return;
    const std::uint64_t after = lane == 63 ? 0 : heads >> (lane + 1);
    const int next = after ? lane + 1 + __builtin_ctzll(after) : 64;
    const std::uint64_t run = (next == 64 ? ~0ull : ((1ull << next) - 1)) & ~((1ull << lane) - 1);
    const int cnt = __popcll(nzm & run);
    if (cnt > 0) atomicAdd(&rnz[r], static_cast<unsigned long long>(cnt));
}

// Per destination row: the longest row, the row-length histogram for the
// apply, and the rows too long for one thread, listed (up to kMidRow entries
// from the front of the list, longer ones from the back).
__global__ void __launch_bounds__(256) row_lengths_kernel(const unsigned long long *rnz, index_t n_dst, int *max_row, std::int32_t *long_rows,
                                                          int *n_long, unsigned long long *len_hist) {
    __shared__ unsigned hist[kHistMax + 2];
    if (threadIdx.x < kHistMax + 2) hist[threadIdx.x] = 0;
    __syncthreads();
    int longest = 0;
    for (index_t r = static_cast<index_t>(blockIdx.x) * 256 + threadIdx.x; r < n_dst; r += static_cast<index_t>(gridDim.x) * 256) {
        const int len = static_cast<int>(min<unsigned long long>(rnz[r], 1ull << 30));
        longest = max(longest, len);
        atomicAdd(&hist[len > kHistMax ? kHistMax + 1 : len], 1u);
        if (len > kShortRow) {
            if (len <= kMidRow)
                long_rows[atomicAdd(&n_long[0], 1)] = static_cast<std::int32_t>(r);
            else
                long_rows[n_dst - 1 - atomicAdd(&n_long[1], 1)] = static_cast<std::int32_t>(r);
        }
    }
    for (int d = 32; d >= 1; d >>= 1) longest = max(longest, __shfl_xor(longest, d, 64));
    __shared__ int part[4];
    if ((threadIdx.x & 63) == 0) part[threadIdx.x / 64] = longest;
    __syncthreads();
    if (threadIdx.x == 0) {
        longest = max(max(part[0], part[1]), max(part[2], part[3]));
        if (longest > 0) atomicMax(max_row, longest);
    }
    if (threadIdx.x < kHistMax + 2 && hist[threadIdx.x]) atomicAdd(&len_hist[threadIdx.x], static_cast<unsigned long long>(hist[threadIdx.x]));
}

// rows of at most kShortRow nonzeros, one thread each: the row's nonzero
// candidates in column order (insertion sort), then weights and frac_b
__global__ void __launch_bounds__(256) short_rows_kernel(const index_t *cand_off, const index_t *row_ptr, index_t n_dst, const std::int32_t *pair_src,
                                                         const double *overlap, const double *area_dst, int frac_area, index_t *col, double *val,
                                                         double *ov_s, double *frac_b) {
    for (index_t r = static_cast<index_t>(blockIdx.x) * 256 + threadIdx.x; r < n_dst; r += static_cast<index_t>(gridDim.x) * 256) {
        const index_t b = row_ptr[r];
        const int len = static_cast<int>(min<index_t>(row_ptr[r + 1] - b, index_t(1) << 30));
        if (len > kShortRow) continue;  // (a wave or a workgroup takes it)
        std::int32_t cc[kShortRow];
        double oo[kShortRow];
        int k = 0;
        // candidates kShortBatch at a time: their overlaps and sources loaded
        // together (the sources whatever the overlap), then filed in order
        const index_t e = cand_off[r + 1];
        for (index_t p0 = cand_off[r]; p0 < e && k < len; p0 += kShortBatch) {
            double ob[kShortBatch];
            std::int32_t sb[kShortBatch];
#pragma unroll
            for (int j = 0; j < kShortBatch; ++j) {
                ob[j] = 0.0;
                sb[j] = 0;
                if (p0 + j < e) {
                    ob[j] = overlap[p0 + j];
                    sb[j] = pair_src[p0 + j];
                }
            }
#pragma unroll
            for (int j = 0; j < kShortBatch; ++j) {
                const double o = ob[j];
                if (!(o > 0.0) || k >= len) continue;
                const std::int32_t c = sb[j];
                int m = k;
                for (; m > 0 && cc[m - 1] > c; --m) {
                    cc[m] = cc[m - 1];
                    oo[m] = oo[m - 1];
                }
                cc[m] = c;
                oo[m] = o;
                ++k;
            }
        }
        const double a = area_dst[r];
        double fb = 0.0;
        for (int q = 0; q < len; ++q) fb += oo[q] / a;
        const bool divide = frac_area && fb > 0.0;
        for (int q = 0; q < len; ++q) {
            double w = fmax(oo[q] / a, 0.0);
            if (divide) w /= fb;
            val[b + q] = w;
            col[b + q] = cc[q];
            ov_s[b + q] = oo[q];
        }
        frac_b[r] = fmin(fb, 1.0);
    }
}

// listed rows of min_len < len <= CAP: bitonic column sort in LDS, the
// ordered sum by one thread, the rest in parallel (columns are distinct
// within a row, so the sorted order does not depend on the sort).  CAP =
// kMidRow with one wave for most long rows (LDS of a few hundred bytes: many
// rows per CU), CAP = kLongRowMax with kLongRowBlock threads for the rest up
// to that length, CAP = kHugeRowMax (one workgroup per CU) beyond.
template <int CAP, int BLOCK>
__global__ void __launch_bounds__(BLOCK) long_rows_kernel(const std::int32_t *long_rows, const int *n_list, bool from_back, int min_len, index_t n_dst,
                                                          const index_t *cand_off, const index_t *row_ptr, const std::int32_t *pair_src,
                                                          const double *overlap, const double *area_dst, int frac_area, index_t *col, double *val,
                                                          double *ov_s, double *frac_b) {
    __shared__ std::int32_t sc[CAP];
    __shared__ double so[CAP];
    __shared__ double s_fb;
    __shared__ int s_n;
    const int tid = static_cast<int>(threadIdx.x);
    const int nl = *n_list;
    for (int li = blockIdx.x; li < nl; li += gridDim.x) {  // (workgroups loop over the listed rows)
        const index_t r = from_back ? long_rows[n_dst - 1 - li] : long_rows[li];
        const index_t b = row_ptr[r];
        const int len = static_cast<int>(row_ptr[r + 1] - b);  // <= the longest listed row (checked by the host)
        if (len <= min_len || len > CAP) continue;           // (another launch takes it; uniform over the workgroup)
        int P = 2;
        while (P < len) P <<= 1;
        // the row's nonzero candidates into LDS, in any order (columns are
        // distinct within a row, so the sort below fixes the order)
        if (tid == 0) s_n = 0;
        __syncthreads();
        for (index_t p = cand_off[r] + tid, e = cand_off[r + 1]; p < e; p += BLOCK) {
            const double o = overlap[p];
            if (o > 0.0) {
                const int k = atomicAdd(&s_n, 1);
                sc[k] = pair_src[p];
                so[k] = o;
            }
        }
        __syncthreads();
        for (int k = len + tid; k < P; k += BLOCK) {
            sc[k] = 0x7FFFFFFF;
            so[k] = 0.0;
        }
        __syncthreads();
        for (int size = 2; size <= P; size <<= 1) {
            for (int stride = size >> 1; stride > 0; stride >>= 1) {
                for (int i = tid; i < P / 2; i += BLOCK) {
                    const int lo = 2 * i - (i & (stride - 1)), hi = lo + stride;
                    const bool asc = (lo & size) == 0;
                    if ((sc[lo] > sc[hi]) == asc) {
                        const std::int32_t tc = sc[lo];
                        sc[lo] = sc[hi];
                        sc[hi] = tc;
                        const double to = so[lo];
                        so[lo] = so[hi];
                        so[hi] = to;
                    }
                }
                __syncthreads();
            }
        }
        // sorted entries out, then the quotients overlap / area in place (in
        // parallel), then their sum in column order by one thread
        const double a = area_dst[r];
        for (int k = tid; k < len; k += BLOCK) {
            col[b + k] = sc[k];
            ov_s[b + k] = so[k];
            so[k] = so[k] / a;
        }
        __syncthreads();
        if (tid == 0) {
            double fb = 0.0;
            for (int k = 0; k < len; ++k) fb += so[k];
            s_fb = fb;
            frac_b[r] = fmin(fb, 1.0);
        }
        __syncthreads();
        const double fb = s_fb;
        const bool divide = frac_area && fb > 0.0;
        for (int k = tid; k < len; k += BLOCK) {
            double w = fmax(so[k], 0.0);
            if (divide) w /= fb;
            val[b + k] = w;
        }
        __syncthreads();  // (the next row reuses the LDS)
    }
}

int bits_for(index_t n) {
    int b = 1;
    while ((index_t(1) << b) < n) ++b;
    return b;
}

// ------------------------------------------------------------ host replay
struct HostSide {
    const Mesh *m;
    const MeshGeom *g;
    // node xyz with host libm: the host array, or (device-built xyz) the node's own
    V3 node(index_t nd) const {
        if (g->dev_xyz) {
            double p[3];
            node_xyz_one(*m, nd, p);
            return V3{p[0], p[1], p[2]};
        }
        return V3{g->h_xyz[3 * nd], g->h_xyz[3 * nd + 1], g->h_xyz[3 * nd + 2]};
    }
    V3 vertex(index_t c, int k) const {
        if (g->regular) {
            const index_t nl = g->ni + 1, na = g->nj + 1;
            const double *slon = g->h_trig.data(), *clon = slon + nl, *slat = clon + nl, *clat = slat + na;
            index_t ci = c % g->ni, cj = c / g->ni;
            index_t a = (k == 1 || k == 2) ? ci + 1 : ci;
            index_t b = (k >= 2) ? cj + 1 : cj;
            return V3{clat[b] * clon[a], clat[b] * slon[a], slat[b]};
        }
        return node(m->h_indices[m->h_offsets[c] + k]);
    }
    int nverts(index_t c) const {
        if (g->regular) return 4;
        return static_cast<int>(std::min<index_t>(m->h_offsets[c + 1] - m->h_offsets[c], 32));
    }
    // the cell's nodes, each evaluated once (node() may call host libm);
    // false for a cell of more than kMaxNodes nodes
    static constexpr int kMaxNodes = 32;
    bool nodes(index_t c, V3 *out, int &n) const {
        const index_t b = m->h_offsets[c], e = m->h_offsets[c + 1];
        if (e - b > kMaxNodes) return false;
        n = static_cast<int>(e - b);
        for (int k = 0; k < n; ++k) out[k] = node(m->h_indices[b + k]);
        return true;
    }
    // precompute_cap_data for one cell with host libm (spherical_cap_filter.hpp:85-167),
    // over the cell's nodes p[0 .. n)
    static void cap_of(const V3 *p, int n, V3 &cen, double &rad) {
        cen = V3{0, 0, 0};
        rad = 0.0;
        if (n < 1) return;
        double sx = 0, sy = 0, sz = 0;
        for (int k = 0; k < n; ++k) {
            sx += p[k].x;
            sy += p[k].y;
            sz += p[k].z;
        }
        double inv = 1.0 / static_cast<double>(n);
        cen = geo::normalize(V3{sx * inv, sy * inv, sz * inv});
        for (int k = 0; k < n; ++k) {
            double d = geo::dot(cen, p[k]);
            d = std::fmin(1.0, std::fmax(-1.0, d));
            double ang = std::acos(d);
            if (ang > rad) rad = ang;
        }
    }
    // the same for any cell size, node by node
    void cap(index_t c, V3 &cen, double &rad) const {
        const index_t b = m->h_offsets[c], e = m->h_offsets[c + 1];
        cen = V3{0, 0, 0};
        rad = 0.0;
        if (e - b < 1) return;
        double sx = 0, sy = 0, sz = 0;
        for (index_t k = b; k < e; ++k) {
            const V3 p = node(m->h_indices[k]);
            sx += p.x;
            sy += p.y;
            sz += p.z;
        }
        double inv = 1.0 / static_cast<double>(e - b);
        cen = geo::normalize(V3{sx * inv, sy * inv, sz * inv});
        for (index_t k = b; k < e; ++k) {
            double d = geo::dot(cen, node(m->h_indices[k]));
            d = std::fmin(1.0, std::fmax(-1.0, d));
            double ang = std::acos(d);
            if (ang > rad) rad = ang;
        }
    }
};

// The reference's inner-loop body for one (dst j, src i) pair with host
// libm: cap test, clip with SphericalPolygon<32> capacity, fan area.
double host_pair_overlap(const HostSide &S, const HostSide &D, index_t j, index_t i) {
    // each node's xyz once per pair (the same values node() gives each call)
    V3 sn[HostSide::kMaxNodes], dn[HostSide::kMaxNodes];
    int n_sn = 0, n_dn = 0;
    const bool s_nodes = S.nodes(i, sn, n_sn), d_nodes = D.nodes(j, dn, n_dn);
    V3 cs, cd;
    double rs, rd;
    if (s_nodes)
        HostSide::cap_of(sn, n_sn, cs, rs);
    else
        S.cap(i, cs, rs);
    if (d_nodes)
        HostSide::cap_of(dn, n_dn, cd, rd);
    else
        D.cap(j, cd, rd);
    double c = geo::dot(cs, cd);
    c = std::fmin(1.0, std::fmax(-1.0, c));
    if (std::acos(c) > (rs + rd)) return 0.0;
    double ax[32], ay[32], az[32], bx[32], by[32], bz[32];
    PolyRef A{ax, ay, az, 1}, B{bx, by, bz, 1};
    // a non-regular cell's vertex k is its node k
    const bool s_cached = s_nodes && !S.g->regular, d_cached = d_nodes && !D.g->regular;
    const int ns = S.nverts(i);
    for (int k = 0; k < ns; ++k) A.set(k, s_cached ? sn[k] : S.vertex(i, k));
    const int nc = D.nverts(j);
    bool overflow = false;
    auto clipv = [&](int k) { return d_cached ? dn[k] : D.vertex(j, k); };
    int n = geo::clip_polygon(A, ns, B, clipv, nc, 32, &overflow);
    double abs_sum;
    bool zero_num;
    return geo::polygon_area(A, n, &abs_sum, &zero_num);
}

// WeightGenerator::generate's longitude-disjoint check, weight_generator.cpp:1898-1992
// min / max of the cells' mean vertex longitude as order-preserving integer
// keys: grid-stride over the cells, wave then workgroup reduction, one atomic
// per workgroup (per-wave atomics on two addresses serialise in L2)
constexpr int kRangeBlock = 1024;
constexpr index_t kRangeGrid = 256;  // workgroups: one per CU (each ends in an atomic pair on the same two words)

// lows start at the largest key, highs at zero
__global__ void range_init_kernel(unsigned long long *r) {
    if (threadIdx.x < 4) r[threadIdx.x] = (threadIdx.x & 1) ? 0ull : ~0ull;
}

__global__ void __launch_bounds__(kRangeBlock) lon_range_kernel(const index_t *off, const index_t *ind, const double *lon, index_t n_cells,
                                                                unsigned long long *mn, unsigned long long *mx) {
    unsigned long long kmin = ~0ull, kmax = 0ull;
    for (index_t c = static_cast<index_t>(blockIdx.x) * kRangeBlock + threadIdx.x; c < n_cells;
         c += static_cast<index_t>(gridDim.x) * kRangeBlock) {
        index_t b = off[c], e = off[c + 1];
        if (e > b) {
            double s = 0.0;
            if (e - b == 4) {  // a quad: its four indices, then its four longitudes, loaded together
                const index_t i0 = ind[b], i1 = ind[b + 1], i2 = ind[b + 2], i3 = ind[b + 3];
                const double l0 = lon[i0], l1 = lon[i1], l2 = lon[i2], l3 = lon[i3];
                s += l0;
                s += l1;
                s += l2;
                s += l3;
            } else {
                for (index_t k = b; k < e; ++k) s += lon[ind[k]];
            }
            double cx = s / static_cast<double>(e - b);
            std::uint64_t bits = static_cast<std::uint64_t>(__double_as_longlong(cx));
            const unsigned long long key = (bits >> 63) ? ~bits : (bits | 0x8000000000000000ull);
            kmin = key < kmin ? key : kmin;
            kmax = key > kmax ? key : kmax;
        }
    }
    for (int d = 32; d >= 1; d >>= 1) {
        const unsigned long long a = __shfl_xor(kmin, d, 64), z = __shfl_xor(kmax, d, 64);
        kmin = a < kmin ? a : kmin;
        kmax = z > kmax ? z : kmax;
    }
    __shared__ unsigned long long part[2][kRangeBlock / 64];
    if ((threadIdx.x & 63) == 0) {
        part[0][threadIdx.x / 64] = kmin;
        part[1][threadIdx.x / 64] = kmax;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < kRangeBlock / 64; ++w) {
            kmin = part[0][w] < kmin ? part[0][w] : kmin;
            kmax = part[1][w] > kmax ? part[1][w] : kmax;
        }
        if (kmin != ~0ull) atomicMin(mn, kmin);
        if (kmax != 0ull) atomicMax(mx, kmax);
    }
}

double key_to_double(unsigned long long k) {
    std::uint64_t b = (k >> 63) ? (k & 0x7FFFFFFFFFFFFFFFull) : ~static_cast<std::uint64_t>(k);
    double d;
    std::memcpy(&d, &b, sizeof d);
    return d;
}

// longitude ranges of the vertices of both meshes' cells: one pair of
// reductions into r[4] (an empty mesh leaves lo = all ones, hi = 0)
void lon_ranges_launch(const Mesh &a, const Mesh &b, unsigned long long *r, hipStream_t s) {
    range_init_kernel<<<1, 64, 0, s>>>(r);
    AXIS_LAUNCH_CHECK();
    const Mesh *ms[2] = {&a, &b};
    for (int k = 0; k < 2; ++k) {
        const Mesh &m = *ms[k];
        if (m.n_cells() == 0) continue;
        const unsigned grid = static_cast<unsigned>(std::min<index_t>((m.n_cells() + kRangeBlock - 1) / kRangeBlock, kRangeGrid));
        lon_range_kernel<<<grid, kRangeBlock, 0, s>>>(m.d_offsets.get(), m.d_indices.get(), m.d_coords.get(), m.n_cells(), r + 2 * k, r + 2 * k + 1);
        AXIS_LAUNCH_CHECK();
    }
}

void lon_ranges_decode(const unsigned long long h[4], double range[4]) {
    for (int k = 0; k < 2; ++k) {
        if (h[2 * k] == ~0ull) {
            range[2 * k] = std::numeric_limits<double>::max();
            range[2 * k + 1] = -std::numeric_limits<double>::max();
        } else {
            range[2 * k] = key_to_double(h[2 * k]);
            range[2 * k + 1] = key_to_double(h[2 * k + 1]);
        }
    }
}

void lon_ranges(const Mesh &a, const Mesh &b, double range[4], hipStream_t s) {
    DeviceBuffer<unsigned long long> r(4);
    lon_ranges_launch(a, b, r.get(), s);
    unsigned long long h[4];
    read_back(h, r.get(), 4, s);
    lon_ranges_decode(h, range);
}

// the reference's decision on the two ranges (weight_generator.cpp:1898-1992)
void lon_mismatch_check(CoordinateSystem cs, const double range[4], bool structured_fast_path) {
    const double s0 = range[0], s1 = range[1], d0 = range[2], d1 = range[3];
    bool disjoint = (s1 < d0) || (d1 < s0);
    if (!disjoint) return;
    const double period = (cs == CoordinateSystem::SphericalDeg) ? 360.0 : (2.0 * kPi);
    double shift = std::round((0.5 * (s0 + s1) - 0.5 * (d0 + d1)) / period) * period;
    bool overlaps = !(s1 < d0 + shift || d1 + shift < s0);
    if (overlaps && !structured_fast_path)
        throw InvalidArgument("WeightGenerator::generate: Spherical longitude coordinate range mismatch: source [" + std::to_string(s0) + ", " +
                              std::to_string(s1) + "] and destination [" + std::to_string(d0) + ", " + std::to_string(d1) +
                              "] are disjoint but would overlap after a shift of " + std::to_string(shift));
}

}  // namespace

void check_longitude_ranges(const Mesh &src, const Mesh &dst, bool structured_fast_path, hipStream_t s) {
    const auto cs = src.coord_system();
    if (cs != CoordinateSystem::SphericalDeg && cs != CoordinateSystem::SphericalRad) return;
    double range[4];
    lon_ranges(src, dst, range, s);
    lon_mismatch_check(cs, range, structured_fast_path);
}

void LonCheck::launch(const Mesh &src, const Mesh &dst, hipStream_t s) {
    const auto cs = src.coord_system();
    armed_ = cs == CoordinateSystem::SphericalDeg || cs == CoordinateSystem::SphericalRad;
    if (!armed_) return;
    src_ = &src;
    dst_ = &dst;
    d_.resize(4);
    h_.resize(4);
    if (!done_) AXIS_HIP_CHECK(hipEventCreateWithFlags(&done_, hipEventDisableTiming));
    lon_ranges_launch(src, dst, d_.get(), s);
    AXIS_HIP_CHECK(hipMemcpyAsync(h_.data(), d_.get(), 4 * sizeof(unsigned long long), hipMemcpyDeviceToHost, s));
    AXIS_HIP_CHECK(hipEventRecord(done_, s));
}

void LonCheck::finish(bool structured_fast_path) {
    if (!armed_) return;
    armed_ = false;
    AXIS_HIP_CHECK(hipEventSynchronize(done_));
    double range[4];
    lon_ranges_decode(h_.data(), range);
    lon_mismatch_check(src_->coord_system(), range, structured_fast_path);
}

LonCheck::~LonCheck() {
    if (done_) {
        (void)hipEventSynchronize(done_);  // (the readback lands in a block the cache may hand out again)
        (void)hipEventDestroy(done_);
    }
}

namespace {
// HIP events timing one generation's phases on its stream, created once per
// thread and device and read after the generation's last synchronisation
struct PhaseEvents {
    static constexpr int kN = 8;
    hipEvent_t e[kN] = {};
    bool on[kN] = {};
    static PhaseEvents &get() {
        static thread_local std::map<int, PhaseEvents> per_device;
        int dev = 0;
        AXIS_HIP_CHECK(hipGetDevice(&dev));
        PhaseEvents &p = per_device[dev];
        if (!p.e[0])
            for (hipEvent_t &x : p.e) AXIS_HIP_CHECK(hipEventCreate(&x));
        for (bool &b : p.on) b = false;
        return p;
    }
    void rec(int k, hipStream_t s) {
        AXIS_HIP_CHECK(hipEventRecord(e[k], s));
        on[k] = true;
    }
    double ms(int a, int b) const {
        float t = -1.0f;
        if (!on[a] || !on[b] || hipEventElapsedTime(&t, e[a], e[b]) != hipSuccess) return -1.0;
        return t;
    }
};
}  // namespace

std::unique_ptr<InterpolationMatrix> generate_conservative(const Mesh &src, const Mesh &dst, const RegridConfig &cfg, hipStream_t s,
                                                           const mlp::Model *model, ProposerStats *pst, double area_tol, LonCheck *lon) {
    const auto t0 = std::chrono::steady_clock::now();
    PhaseTrace tr("conservative", s);
    PhaseEvents &ev = PhaseEvents::get();
    // the host node-xyz task uploads on its own stream: this thread's
    // allocations must not wait for the device (and thereby for the task);
    // blocks released before this call are made reusable first
    BlockCache::recycle_now();
    BlockCache::NoWaitScope no_wait;
    const index_t n_src = src.n_cells(), n_dst = dst.n_cells();
    if (n_src >= (index_t(1) << 31) || n_dst >= (index_t(1) << 31)) throw InvalidArgument("generate_conservative: more than 2^31 cells");
    MeshGeom sg, dg;
    // host libm node xyz (source, then destination) on a host thread beside
    // the GPU grid detection and the source's upload and cell pass
    // (a structured pair takes an analytic path instead: the task is told to stop)
    // A rectilinear degree source has its node xyz built on the device from
    // axis tables instead.  When the source's counts allow a rectilinear grid
    // the task starts on the destination at once (a source that then turns
    // out to need the host is done on this thread); otherwise it does the
    // source first, and is told to skip it should its axes allow the device.
    std::atomic<bool> cancel{false}, skip_src{false};
    std::promise<void> src_done;
    std::future<void> src_ready = src_done.get_future();
    const bool spherical = src.coord_system() != CoordinateSystem::Cartesian3D;
    const bool dst_first = spherical && maybe_rectilinear_deg(src);
    // the destination's xyz goes to the device slice by slice as the host
    // threads finish them, on a side stream that this stream then waits for
    struct EventGuard {
        hipEvent_t e = nullptr;
        ~EventGuard() {
            if (e) (void)hipEventDestroy(e);
        }
    } dst_up, src_up;
    hipStream_t up = nullptr;
    std::atomic<bool> dst_uploaded{false}, src_uploaded{false};
    std::atomic<bool> dst_trig{false}, src_trig{false};  // built by node_xyz_trig (else host libm slices)
    // both meshes with grid counts: an analytic path will most likely answer,
    // so no host task is started (should both decline, the geometry below
    // computes the node xyz on this thread)
    const bool likely_structured = dst_first && maybe_rectilinear_deg(dst);
    std::future<void> xyz_task;
    if (spherical && !likely_structured) {
        up = upload_stream();
        AXIS_HIP_CHECK(hipEventCreateWithFlags(&dst_up.e, hipEventDisableTiming));
        dg.d_xyz.resize(std::max<std::size_t>(3 * static_cast<std::size_t>(dst.n_nodes()), 3));
        double *src_out = nullptr, *dst_out = dg.d_xyz.get();  // (the task never touches the DeviceBuffer objects)
        if (!dst_first) {  // the source's slices go up the same way
            AXIS_HIP_CHECK(hipEventCreateWithFlags(&src_up.e, hipEventDisableTiming));
            sg.d_xyz.resize(std::max<std::size_t>(3 * static_cast<std::size_t>(src.n_nodes()), 3));
            src_out = sg.d_xyz.get();
        }
        xyz_task = std::async(std::launch::async, [&, src_out, dst_out] {
            BlockCache::NoWaitScope task_no_wait;  // (its pinned blocks)
            // on the device where the mesh allows, else host libm by slices
            auto build = [&](const Mesh &m, MeshGeom &g, const std::atomic<bool> *stop, hipEvent_t done, double *out, std::atomic<bool> &trig) {
                if (node_xyz_trig(m, out, up, stop)) {
                    trig.store(true);
                    AXIS_HIP_CHECK(hipEventRecord(done, up));
                    return true;
                }
                return node_xyz_host(m, g, stop, up, done, out);
            };
            try {
                if (!dst_first) src_uploaded.store(build(src, sg, &skip_src, src_up.e, src_out, src_trig));
                src_done.set_value();
            } catch (...) {
                src_done.set_exception(std::current_exception());
                return;
            }
            dst_uploaded.store(build(dst, dg, &cancel, dst_up.e, dst_out, dst_trig));
        });
    } else {
        src_done.set_value();
    }
    GridAxes sax = unique_axes(src, s);
    if (lon) lon->finish(false);  // (its readback was queued before the detection's)
    const bool src_dev_xyz = spherical && node_xyz_device_eligible(src, sax);
    if (src_dev_xyz) skip_src.store(true);
    // the source's axis sin/cos tables (host libm, ~2,000 angles at 0.25deg)
    // on a helper thread while the destination's detection runs
    std::future<PinnedArray<double>> src_tables;
    if (src_dev_xyz)
        src_tables = std::async(std::launch::async, [&sax] {
            BlockCache::NoWaitScope no_wait_here;  // (its pinned block)
            return axis_trig_tables(sax);
        });
    GridAxes dax = unique_axes(dst, s);
    const bool s_reg = regular_from_axes(src, sax, sg), d_reg = regular_from_axes(dst, dax, dg);
    // generate_conservative's dispatch (weight_generator.cpp:3259-3296): the
    // regular, then the rectilinear, analytic path when both meshes qualify;
    // a declined path, or an empty result under UnmappedAction::Error, falls
    // through.
    auto accept = [&](const std::unique_ptr<InterpolationMatrix> &R) {
        return R && n_dst > 0 && (cfg.unmapped != UnmappedAction::Error || R->nnz > 0);
    };
    std::unique_ptr<InterpolationMatrix> fast;
    if (s_reg && d_reg) {
        fast = conservative_regular(src, dst, cfg, sax, dax, sg, dg, s);
        if (!accept(fast)) fast.reset();
    }
    if (!fast && rectilinear_from_axes(src, sax) && rectilinear_from_axes(dst, dax)) {
        fast = conservative_rectilinear(src, dst, cfg, sax, dax, s);
        if (!accept(fast)) fast.reset();
    }
    if (fast) {
        skip_src.store(true);
        cancel.store(true);
        if (xyz_task.valid()) xyz_task.get();
        return fast;
    }
    if (!spherical) throw InvalidArgument("generate_conservative: Cartesian3D meshes are served by the rectangle paths only here");
    if (cfg.line_type != LineType::GreatCircle) throw InvalidArgument("generate_conservative: only LineType::GreatCircle is supported");
    tr.mark("detect");
    // the host task is done with the source (finished, or stopped by
    // skip_src) before the device build may replace sg.d_xyz, so no slice
    // upload is queued into a buffer after its release
    src_ready.get();
    if (src_dev_xyz) {
        PinnedArray<double> tables = src_tables.get();
        node_xyz_device(src, sax, sg, s, &tables);  // (false: host xyz below)
    }
    if (!sg.dev_xyz && src_uploaded.load()) {
        AXIS_HIP_CHECK(hipStreamWaitEvent(s, src_up.e, 0));
        (src_trig.load() ? sg.dev_xyz : sg.d_xyz_ready) = true;
    }
    prepare_geometry(src, sg, s);
    tr.mark("geometry_src");
    // broad phase: the cell grid of a rectilinear degree source, else a BVH
    // (the proposer's fallback rows query the BVH too); neither needs the
    // destination, so it is built while the destination's geometry finishes
    GridIndex grid;
    const char *bp = std::getenv("AXIS_BROAD_PHASE");
    const bool grid_ok = !model && !(bp && std::string(bp) == "lbvh") && rectilinear_from_axes(src, sax) &&
                         grid_index_build(src, sax, sg.d_box.get(), grid, s);
    LbvhTree tree;
    if (!grid_ok) lbvh_build(sg.d_box.get(), n_src, tree, s);
    tr.mark("broad_phase_build");
    if (xyz_task.valid()) xyz_task.get();
    if (dst_uploaded.load()) {
        AXIS_HIP_CHECK(hipStreamWaitEvent(s, dst_up.e, 0));
        (dst_trig.load() ? dg.dev_xyz : dg.d_xyz_ready) = true;
    }
    prepare_geometry(dst, dg, s);
    tr.mark("geometry_dst");

    DeviceBuffer<index_t> cand_off;
    DeviceBuffer<std::int32_t> cand_src, accepted;
    DeviceBuffer<std::int32_t> grid_rows;  // (the grid query writes each candidate's row)
    index_t n_pairs = 0;
    ev.rec(0, s);
    if (grid_ok) {
        n_pairs = grid_query(grid, dg.d_box.get(), n_dst, cand_off, cand_src, grid_rows, s);
    } else {
        if (model)
            proposer_candidates(sg, dg, *model, cand_off, cand_src, accepted, s);
        else
            lbvh_query(tree, dg.d_box.get(), n_dst, cand_off, cand_src, s);
        read_back(&n_pairs, cand_off.get() + n_dst, 1, s);
    }
    ev.rec(1, s);
    tr.mark(model ? "propose" : "broad_phase");

    auto M = std::make_unique<InterpolationMatrix>();
    M->n_src = n_src;
    M->n_dst = n_dst;
    cell_morton_keys(dst, M->row_key, s);

    // with a proposer, room for the LBVH pairs of redone rows (~10%) so that
    // round 2 usually appends in place
    const index_t cap_pairs = model ? n_pairs + n_pairs / 8 + 4096 : std::max<index_t>(n_pairs, 1);
    DeviceBuffer<std::int32_t> pair_row;
    if (grid_ok)
        pair_row = std::move(grid_rows);
    else
        pair_row.resize(cap_pairs);
    DeviceBuffer<double> overlap(cap_pairs);
    DeviceBuffer<std::int32_t> arb(cap_pairs);
    if (model && cand_src.size() < static_cast<std::size_t>(cap_pairs)) {
        DeviceBuffer<std::int32_t> cs(cap_pairs);
        if (n_pairs > 0) AXIS_HIP_CHECK(hipMemcpyAsync(cs.get(), cand_src.get(), sizeof(std::int32_t) * n_pairs, hipMemcpyDeviceToDevice, s));
        cand_src = std::move(cs);
    }
    DeviceBuffer<int> counters(2);
    DeviceBuffer<std::int32_t> trip(3 * static_cast<std::size_t>(kTripCap));
    AXIS_HIP_CHECK(hipMemsetAsync(counters.get(), 0, 2 * sizeof(int), s));
    if (n_dst > 0 && !grid_ok) {
        fill_rows_kernel<<<grid_for(n_dst, 256), 256, 0, s>>>(cand_off.get(), n_dst, pair_row.get());
        AXIS_LAUNCH_CHECK();
    }
    auto side = [](const Mesh &m, const MeshGeom &g) {
        Side x{};
        x.off = m.d_offsets.get();
        x.ind = m.d_indices.get();
        x.xyz = g.d_xyz.get();
        x.trig = g.regular ? g.d_trig.get() : nullptr;
        x.ni = g.ni;
        x.nlon = g.ni + 1;
        x.nlat = g.nj + 1;
        x.flags = g.d_flags.get();
        x.cap = g.d_cap.get();
        x.area = g.d_area.get();
        return x;
    };
    const Side S = side(src, sg), D = side(dst, dg);
    const int nv_s = sg.regular ? 4 : std::min(src.max_cell_vertices(), 32);
    const int nv_d = dg.regular ? 4 : std::min(dst.max_cell_vertices(), 32);
    const int need = nv_s + nv_d;  // bound for convex cells; larger results go to the host
    // overlaps of pairs [p0, p1); pairs the device cannot decide are queued in arb
    auto run_overlaps = [&](index_t p0, index_t p1) {
        if (p1 <= p0) return;
        const unsigned grid = static_cast<unsigned>((p1 - p0 + kWave - 1) / kWave);
        if (p1 - p0 < (index_t(1) << 31)) {
            // cap tests and per-chunk compaction, then the clipping of the
            // survivors; clipping blocks past their chunk's count return at once
            const index_t n_chunks = (p1 - p0 + kSelChunk - 1) / kSelChunk;
            if (n_chunks * (kSelChunk / kWave) >= (index_t(1) << 31)) throw InvalidArgument("generate: too many candidate pairs");
            DeviceBuffer<std::uint32_t> live(n_chunks * kSelChunk);
            DeviceBuffer<std::int32_t> n_live(n_chunks);
            if (p0 == 0) ev.rec(2, s);
            cap_select_kernel<<<static_cast<unsigned>(n_chunks), kSelThreads, 0, s>>>(p0, p1, pair_row.get(), cand_src.get(), S, D, overlap.get(),
                                                                                     live.get(), n_live.get(), counters.get() + 1);
            AXIS_LAUNCH_CHECK();
            if (p0 == 0) ev.rec(3, s);
            const unsigned cgrid = static_cast<unsigned>(n_chunks * (kSelChunk / kWave));
            if (need <= 8)
                clip_live_kernel<8><<<cgrid, kWave, 0, s>>>(p0, live.get(), n_live.get(), static_cast<unsigned>(n_chunks), pair_row.get(), cand_src.get(), S, D, overlap.get(),
                                                            arb.get(), counters.get(), trip.get());
            else if (need <= 16)
                clip_live_kernel<16><<<cgrid, kWave, 0, s>>>(p0, live.get(), n_live.get(), static_cast<unsigned>(n_chunks), pair_row.get(), cand_src.get(), S, D, overlap.get(),
                                                             arb.get(), counters.get(), trip.get());
            else
                clip_live_kernel<24><<<cgrid, kWave, 0, s>>>(p0, live.get(), n_live.get(), static_cast<unsigned>(n_chunks), pair_row.get(), cand_src.get(), S, D, overlap.get(),
                                                             arb.get(), counters.get(), trip.get());
            AXIS_LAUNCH_CHECK();
            if (p0 == 0) ev.rec(4, s);
            return;
        }
        if (need <= 8)
            overlap_kernel<8><<<grid, kWave, 0, s>>>(p0, p1, pair_row.get(), cand_src.get(), S, D, overlap.get(), arb.get(), counters.get(),
                                                     counters.get() + 1, trip.get());
        else if (need <= 16)
            overlap_kernel<16><<<grid, kWave, 0, s>>>(p0, p1, pair_row.get(), cand_src.get(), S, D, overlap.get(), arb.get(), counters.get(),
                                                      counters.get() + 1, trip.get());
        else
            overlap_kernel<24><<<grid, kWave, 0, s>>>(p0, p1, pair_row.get(), cand_src.get(), S, D, overlap.get(), arb.get(), counters.get(),
                                                      counters.get() + 1, trip.get());
        AXIS_LAUNCH_CHECK();
    };
    // host libm decides the queued pairs arb[a0 .. current count)
    int arb_done = 0;
    auto arbitrate = [&]() {
        // the counters and the first queued triples in one readback
        constexpr int kFetch = 4096;
        const int fetch = std::max(0, std::min(kFetch, kTripCap - arb_done));
        PinnedArray<std::int32_t> h;
        h.resize(2 + 3 * static_cast<std::size_t>(kFetch));
        AXIS_HIP_CHECK(hipMemcpyAsync(h.data(), counters.get(), 2 * sizeof(int), hipMemcpyDeviceToHost, s));
        if (fetch > 0)
            AXIS_HIP_CHECK(hipMemcpyAsync(h.data() + 2, trip.get() + 3 * static_cast<std::size_t>(arb_done), 3 * sizeof(std::int32_t) * fetch,
                                          hipMemcpyDeviceToHost, s));
        AXIS_HIP_CHECK(hipStreamSynchronize(s));
        M->n_host_arbitrated = h[0];
        M->n_cap_borderline = h[1];
        const int na = h[0] - arb_done;
        if (na <= 0) return;
        const std::int32_t *tr3 = h.data() + 2;
        std::vector<std::int32_t> more;  // beyond the first fetch
        if (na > fetch) {
            more.resize(3 * static_cast<std::size_t>(na));
            if (arb_done + na <= kTripCap) {
                AXIS_HIP_CHECK(hipMemcpyAsync(more.data(), trip.get() + 3 * static_cast<std::size_t>(arb_done), 3 * sizeof(std::int32_t) * na,
                                              hipMemcpyDeviceToHost, s));
            } else {  // (more than kTripCap queued: gathered from arb)
                DeviceBuffer<std::int32_t> d_trip(3 * static_cast<std::size_t>(na));
                gather_arbitrated_kernel<<<grid_for(na, 256), 256, 0, s>>>(arb.get() + arb_done, na, pair_row.get(), cand_src.get(), d_trip.get());
                AXIS_LAUNCH_CHECK();
                d_trip.download(more.data(), 3 * static_cast<std::size_t>(na), s);
            }
            AXIS_HIP_CHECK(hipStreamSynchronize(s));
            tr3 = more.data();
        }
        PinnedArray<std::int32_t> pairs;
        PinnedArray<double> vals;
        pairs.resize(na);
        vals.resize(na);
        HostSide HS{&src, &sg}, HD{&dst, &dg};
        parallel_range(
            na, [&](index_t b, index_t e) {
                for (index_t k = b; k < e; ++k) {
                    pairs[k] = tr3[3 * k];
                    vals[k] = host_pair_overlap(HS, HD, tr3[3 * k + 1], tr3[3 * k + 2]);
                }
            },
            4);
        // scattered in stream order (the pinned arrays are reused only after
        // the next device synchronisation)
        DeviceBuffer<std::int32_t> d_idx(na);
        DeviceBuffer<double> d_vals(na);
        AXIS_HIP_CHECK(hipMemcpyAsync(d_idx.get(), pairs.data(), sizeof(std::int32_t) * na, hipMemcpyHostToDevice, s));
        AXIS_HIP_CHECK(hipMemcpyAsync(d_vals.get(), vals.data(), sizeof(double) * na, hipMemcpyHostToDevice, s));
        scatter_overlaps_kernel<<<grid_for(na, 256), 256, 0, s>>>(d_idx.get(), d_vals.get(), na, overlap.get());
        AXIS_LAUNCH_CHECK();
        arb_done = h[0];
    };
    run_overlaps(0, n_pairs);
    tr.mark("overlap");
    arbitrate();
    index_t n_cand = n_pairs;
    bool appended = false;  // round-2 pairs sit after all of round 1, out of row order
    if (model) {
        // round 2: the LBVH redoes the rows whose proposal was rejected or
        // whose overlaps do not account for the cell's area
        const index_t nd1 = std::max<index_t>(n_dst, 1);
        DeviceBuffer<index_t> redo(nd1), pos(nd1 + 1);
        DeviceBuffer<unsigned long long> nfail(2);
        AXIS_HIP_CHECK(hipMemsetAsync(nfail.get(), 0, 2 * sizeof(unsigned long long), s));
        if (n_dst > 0) {
            redo_rows_kernel<<<grid_for(n_dst, 256), 256, 0, s>>>(cand_off.get(), n_dst, overlap.get(), dg.d_area.get(), dg.d_flags.get(),
                                                                 accepted.get(), area_tol > 0.0 ? area_tol : 1e-12, redo.get(), nfail.get(),
                                                                 nfail.get() + 1);
            AXIS_LAUNCH_CHECK();
        }
        exclusive_scan_i64(redo.get(), pos.get(), n_dst, s);
        index_t n_redo = 0;
        unsigned long long h_fail[2] = {0, 0};
        AXIS_HIP_CHECK(hipMemcpyAsync(&n_redo, pos.get() + n_dst, sizeof(index_t), hipMemcpyDeviceToHost, s));
        AXIS_HIP_CHECK(hipMemcpyAsync(h_fail, nfail.get(), sizeof h_fail, hipMemcpyDeviceToHost, s));
        AXIS_HIP_CHECK(hipStreamSynchronize(s));
        if (pst) {
            pst->rows = n_dst;
            pst->guard_rejected = static_cast<index_t>(h_fail[0]);
            pst->area_failed = static_cast<index_t>(h_fail[1]);
            pst->pairs_proposed = n_pairs;
            pst->pairs_fallback = 0;
        }
        if (n_redo > 0) {
            DeviceBuffer<std::int32_t> rows(n_redo);
            DeviceBuffer<float> rbox(6 * n_redo);
            redo_list_kernel<<<grid_for(n_dst, 256), 256, 0, s>>>(redo.get(), pos.get(), n_dst, dg.d_box.get(), rows.get(), rbox.get());
            AXIS_LAUNCH_CHECK();
            zero_rows_kernel<<<grid_for(n_redo, 256), 256, 0, s>>>(rows.get(), n_redo, cand_off.get(), overlap.get());
            AXIS_LAUNCH_CHECK();
            DeviceBuffer<index_t> off2;
            DeviceBuffer<std::int32_t> src2;
            lbvh_query(tree, rbox.get(), n_redo, off2, src2, s);
            index_t n2 = 0;
            AXIS_HIP_CHECK(hipMemcpyAsync(&n2, off2.get() + n_redo, sizeof(index_t), hipMemcpyDeviceToHost, s));
            AXIS_HIP_CHECK(hipStreamSynchronize(s));
            if (n_pairs + n2 >= (index_t(1) << 31)) throw RuntimeError("generate_conservative: candidate pair count exceeds 2^31");
            const index_t nt = n_pairs + n2;
            if (nt > cap_pairs) {  // grow the pair arrays, keeping round 1
                DeviceBuffer<std::int32_t> pr(nt), cs(nt), ab(nt);
                DeviceBuffer<double> ovb(nt);
                if (n_pairs > 0) {
                    AXIS_HIP_CHECK(hipMemcpyAsync(pr.get(), pair_row.get(), sizeof(std::int32_t) * n_pairs, hipMemcpyDeviceToDevice, s));
                    AXIS_HIP_CHECK(hipMemcpyAsync(cs.get(), cand_src.get(), sizeof(std::int32_t) * n_pairs, hipMemcpyDeviceToDevice, s));
                    AXIS_HIP_CHECK(hipMemcpyAsync(ovb.get(), overlap.get(), sizeof(double) * n_pairs, hipMemcpyDeviceToDevice, s));
                }
                if (arb_done > 0) AXIS_HIP_CHECK(hipMemcpyAsync(ab.get(), arb.get(), sizeof(std::int32_t) * arb_done, hipMemcpyDeviceToDevice, s));
                AXIS_HIP_CHECK(hipStreamSynchronize(s));
                pair_row = std::move(pr);
                cand_src = std::move(cs);
                overlap = std::move(ovb);
                arb = std::move(ab);
            }
            if (n2 > 0) {
                AXIS_HIP_CHECK(hipMemcpyAsync(cand_src.get() + n_pairs, src2.get(), sizeof(std::int32_t) * n2, hipMemcpyDeviceToDevice, s));
                fill_rows_mapped_kernel<<<grid_for(n_redo, 256), 256, 0, s>>>(off2.get(), n_redo, rows.get(), pair_row.get() + n_pairs);
                AXIS_LAUNCH_CHECK();
            }
            appended = n2 > 0;
            run_overlaps(n_pairs, nt);
            arbitrate();
            n_pairs = nt;
            n_cand = nt;
            if (pst) pst->pairs_fallback = n2;
        }
        tr.mark("fallback");
    }
    M->n_candidates = n_cand;

    tr.mark("arbitration");
    // ---- CSR assembly
    ev.rec(5, s);
    const int col_bits = bits_for(std::max<index_t>(n_src, 2));
    const int row_bits = bits_for(std::max<index_t>(n_dst, 2));
    const int frac_area = cfg.norm_type == NormType::FracArea ? 1 : 0;
    constexpr int kH = sizeof(M->len_hist) / sizeof(M->len_hist[0]);
    static_assert(kH == kHistMax + 2, "row-length histogram size");
    M->row_ptr.resize(n_dst + 1);
    M->frac_a.resize(std::max<index_t>(n_src, 1));
    M->frac_b.resize(std::max<index_t>(n_dst, 1));
    DeviceBuffer<int> max_row(1);
    DeviceBuffer<unsigned long long> d_hist(kH);
    AXIS_HIP_CHECK(hipMemsetAsync(max_row.get(), 0, sizeof(int), s));
    AXIS_HIP_CHECK(hipMemsetAsync(d_hist.get(), 0, sizeof(unsigned long long) * kH, s));
    // frac_a: each column's entries gathered by counting and summed in
    // ascending entry (row) order; a column of more than kColLds entries (a
    // source cell under that many destination cells) sets the flag, and the
    // stable global column sort below then gives every column's sum.  Where
    // the pairs are grouped by row with rows ascending, the entries are the
    // candidates of positive overlap, counted on a side stream beside the
    // CSR assembly; otherwise the CSR's entries, after it.
    DeviceBuffer<int> col_flags(2);  // long-column count, overflow
    const char *fa = std::getenv("AXIS_FRAC_A");  // "sort": the global column sort for every column
    const bool fa_sort = fa && std::string(fa) == "sort";
    const bool by_cand = !appended && n_src > 0 && n_pairs < (index_t(1) << 32) && !fa_sort;
    hipStream_t fs = s;
    struct FracEvents {
        hipEvent_t ready = nullptr, done = nullptr;
        ~FracEvents() {
            if (ready) (void)hipEventDestroy(ready);
            if (done) (void)hipEventDestroy(done);
        }
    } fev;
    DeviceBuffer<unsigned long long> col_cnt;
    DeviceBuffer<index_t> col_ptr;
    DeviceBuffer<std::uint32_t> col_ent;
    DeviceBuffer<std::int32_t> long_cols;
    DeviceBuffer<std::uint32_t> col_rank;
    const double *ov_csr = nullptr;  // (the CSR-order overlaps, once they exist)
    auto frac_a_counted = [&](hipStream_t st, bool cand) {
        const index_t n_ent = cand ? n_pairs : M->nnz;
        const double *keep = cand ? overlap.get() : nullptr;
        const double *ov_e = cand ? overlap.get() : ov_csr;
        col_cnt.resize(n_src);
        col_ptr.resize(n_src + 1);
        col_ent.resize(std::max<index_t>(n_ent, 1));
        col_rank.resize(std::max<index_t>(n_ent, 1));
        long_cols.resize(n_src);
        AXIS_HIP_CHECK(hipMemsetAsync(col_cnt.get(), 0, sizeof(unsigned long long) * n_src, st));
        if (n_ent > 0) {
            if (cand)
                col_count_kernel<<<grid_for(n_ent, 256), 256, 0, st>>>(cand_src.get(), keep, n_ent, col_cnt.get(), col_rank.get());
            else
                col_count_kernel<<<grid_for(n_ent, 256), 256, 0, st>>>(M->col.get(), keep, n_ent, col_cnt.get(), col_rank.get());
            AXIS_LAUNCH_CHECK();
        }
        exclusive_scan_i64(reinterpret_cast<const index_t *>(col_cnt.get()), col_ptr.get(), n_src, st);
        if (n_ent > 0) {
            if (cand)
                col_scatter_kernel<<<grid_for(n_ent, 256), 256, 0, st>>>(cand_src.get(), keep, n_ent, col_ptr.get(), col_rank.get(), col_ent.get());
            else
                col_scatter_kernel<<<grid_for(n_ent, 256), 256, 0, st>>>(M->col.get(), keep, n_ent, col_ptr.get(), col_rank.get(), col_ent.get());
            AXIS_LAUNCH_CHECK();
        }
        frac_a_short_kernel<<<grid_for(n_src, kColBlock), kColBlock, 0, st>>>(col_ptr.get(), col_ent.get(), n_src, ov_e, sg.d_area.get(),
                                                                             M->frac_a.get(), long_cols.get(), col_flags.get());
        AXIS_LAUNCH_CHECK();
        frac_a_long_kernel<<<1024, kColBlock, 0, st>>>(col_ptr.get(), col_ent.get(), long_cols.get(), col_flags.get(), ov_e, sg.d_area.get(),
                                                       M->frac_a.get(), col_flags.get() + 1);
        AXIS_LAUNCH_CHECK();
    };
    if (by_cand) {
        fs = upload_stream();
        AXIS_HIP_CHECK(hipEventCreateWithFlags(&fev.ready, hipEventDisableTiming));
        AXIS_HIP_CHECK(hipEventCreateWithFlags(&fev.done, hipEventDisableTiming));
        AXIS_HIP_CHECK(hipEventRecord(fev.ready, s));
        AXIS_HIP_CHECK(hipStreamWaitEvent(fs, fev.ready, 0));
        AXIS_HIP_CHECK(hipMemsetAsync(col_flags.get(), 0, 2 * sizeof(int), fs));
        frac_a_counted(fs, true);
        AXIS_HIP_CHECK(hipEventRecord(fev.done, fs));
    } else {
        AXIS_HIP_CHECK(hipMemsetAsync(col_flags.get(), 0, 2 * sizeof(int), s));
    }
    // pairs grouped by row (every path but a proposer round 2 that appended
    // pairs): per-row nonzero counts give row_ptr, then each row is sorted by
    // column where it lies (one thread, one wave or one workgroup per row)
    DeviceBuffer<std::int32_t> long_rows;
    DeviceBuffer<int> n_long(2);  // rows of up to kMidRow entries, longer rows
    index_t nnz = 0;
    int h_max_row = 0;
    bool grouped = !appended && n_dst > 0;
    {
        PinnedArray<index_t> h;
        h.resize(2);
        h[0] = 0;
        h[1] = 0;
        if (grouped) {
            long_rows.resize(n_dst);
            DeviceBuffer<unsigned long long> rnz(n_dst);
            AXIS_HIP_CHECK(hipMemsetAsync(n_long.get(), 0, 2 * sizeof(int), s));
            AXIS_HIP_CHECK(hipMemsetAsync(rnz.get(), 0, sizeof(unsigned long long) * n_dst, s));
            if (n_pairs > 0) {
                nz_count_kernel<<<grid_for(n_pairs, 256), 256, 0, s>>>(overlap.get(), pair_row.get(), n_pairs, rnz.get());
                AXIS_LAUNCH_CHECK();
            }
            const unsigned grid = static_cast<unsigned>(std::min<index_t>((n_dst + 255) / 256, 2048));
            row_lengths_kernel<<<grid, 256, 0, s>>>(rnz.get(), n_dst, max_row.get(), long_rows.get(), n_long.get(), d_hist.get());
            AXIS_LAUNCH_CHECK();
            exclusive_scan_i64(reinterpret_cast<const index_t *>(rnz.get()), M->row_ptr.get(), n_dst, s);
            AXIS_HIP_CHECK(hipMemcpyAsync(h.data(), M->row_ptr.get() + n_dst, sizeof(index_t), hipMemcpyDeviceToHost, s));
            AXIS_HIP_CHECK(hipMemcpyAsync(h.data() + 1, max_row.get(), sizeof(int), hipMemcpyDeviceToHost, s));
            AXIS_HIP_CHECK(hipStreamSynchronize(s));
            nnz = h[0];
            h_max_row = static_cast<int>(h[1] & 0xffffffff);
            grouped = h_max_row <= kHugeRowMax;
        }
    }
    DeviceBuffer<index_t> pos;
    if (!grouped) {  // nonzero flags over the pairs and their scan (the global sort below)
        DeviceBuffer<index_t> nzf(std::max<index_t>(n_pairs, 1));
        pos.resize(n_pairs + 1);
        if (n_pairs > 0) {
            nz_flag_kernel<<<grid_for(n_pairs, 256), 256, 0, s>>>(overlap.get(), n_pairs, nzf.get());
            AXIS_LAUNCH_CHECK();
        }
        exclusive_scan_i64(nzf.get(), pos.get(), n_pairs, s);
        read_back(&nnz, pos.get() + n_pairs, 1, s);
    }
    M->nnz = nnz;
    const index_t nz1 = std::max<index_t>(nnz, 1);
    M->col.resize(nz1);
    M->val.resize(nz1);
    DeviceBuffer<double> ov(nz1);  // overlaps in CSR order
    DeviceBuffer<std::uint32_t> perm(nz1);
    bool assembled = false;
    if (grouped) {
        // the listed rows on a second stream, beside the short rows (they
        // write other rows); this stream waits for them
        hipStream_t ls = s;
        FracEvents lev;
        if (h_max_row > kShortRow) {
            ls = rows_stream();
            AXIS_HIP_CHECK(hipEventCreateWithFlags(&lev.ready, hipEventDisableTiming));
            AXIS_HIP_CHECK(hipEventCreateWithFlags(&lev.done, hipEventDisableTiming));
            AXIS_HIP_CHECK(hipEventRecord(lev.ready, s));
            AXIS_HIP_CHECK(hipStreamWaitEvent(ls, lev.ready, 0));
        }
        const unsigned grid = static_cast<unsigned>(std::min<index_t>((n_dst + 255) / 256, 2048));
        short_rows_kernel<<<grid, 256, 0, s>>>(cand_off.get(), M->row_ptr.get(), n_dst, cand_src.get(), overlap.get(), dg.d_area.get(), frac_area,
                                               M->col.get(), M->val.get(), ov.get(), M->frac_b.get());
        AXIS_LAUNCH_CHECK();
        if (h_max_row > kShortRow) {  // the listed rows (their counts stay on the device)
            long_rows_kernel<kMidRow, kMidRow><<<2048, kMidRow, 0, ls>>>(long_rows.get(), n_long.get(), false, kShortRow, n_dst, cand_off.get(),
                                                                        M->row_ptr.get(), cand_src.get(), overlap.get(), dg.d_area.get(),
                                                                        frac_area, M->col.get(), M->val.get(), ov.get(), M->frac_b.get());
            AXIS_LAUNCH_CHECK();
        }
        if (h_max_row > kMidRow) {
            long_rows_kernel<kLongRowMax, kLongRowBlock><<<512, kLongRowBlock, 0, ls>>>(
                long_rows.get(), n_long.get() + 1, true, kMidRow, n_dst, cand_off.get(), M->row_ptr.get(), cand_src.get(), overlap.get(),
                dg.d_area.get(), frac_area, M->col.get(), M->val.get(), ov.get(), M->frac_b.get());
            AXIS_LAUNCH_CHECK();
        }
        if (h_max_row > kLongRowMax) {
            long_rows_kernel<kHugeRowMax, kHugeRowBlock><<<256, kHugeRowBlock, 0, ls>>>(
                long_rows.get(), n_long.get() + 1, true, kLongRowMax, n_dst, cand_off.get(), M->row_ptr.get(), cand_src.get(), overlap.get(),
                dg.d_area.get(), frac_area, M->col.get(), M->val.get(), ov.get(), M->frac_b.get());
            AXIS_LAUNCH_CHECK();
        }
        if (ls != s) {
            AXIS_HIP_CHECK(hipEventRecord(lev.done, ls));
            AXIS_HIP_CHECK(hipStreamWaitEvent(s, lev.done, 0));
        }
        assembled = true;
    }
    if (!assembled) {
        // pairs in any order (or rows beyond the LDS sort): one global (row, col) sort
        DeviceBuffer<std::uint64_t> key(nz1);
        DeviceBuffer<double> ov_c(nz1);
        if (nnz > 0) {
            compact_kernel<<<grid_for(n_pairs, 256), 256, 0, s>>>(overlap.get(), pos.get(), n_pairs, pair_row.get(), cand_src.get(), col_bits,
                                                                 key.get(), perm.get(), ov_c.get());
            AXIS_LAUNCH_CHECK();
            radix_sort_pairs_u64(key.get(), perm.get(), nnz, col_bits + row_bits, s);
            unpack_sorted_kernel<<<grid_for(nnz, 256), 256, 0, s>>>(key.get(), perm.get(), ov_c.get(), nnz, col_bits, n_dst, M->col.get(), ov.get(),
                                                                    M->row_ptr.get());
            AXIS_LAUNCH_CHECK();
        } else {
            AXIS_HIP_CHECK(hipMemsetAsync(M->row_ptr.get(), 0, sizeof(index_t) * (n_dst + 1), s));
        }
        AXIS_HIP_CHECK(hipMemsetAsync(max_row.get(), 0, sizeof(int), s));
        if (n_dst > 0) {
            const unsigned grid = static_cast<unsigned>(std::min<index_t>((n_dst + 255) / 256, 2048));
            row_weights_kernel<<<grid, 256, 0, s>>>(M->row_ptr.get(), n_dst, dg.d_area.get(), ov.get(), M->val.get(), M->frac_b.get(), frac_area,
                                                    max_row.get());
            AXIS_LAUNCH_CHECK();
        }
    }
    ov_csr = ov.get();
    const bool by_count = by_cand || (n_src > 0 && nnz < (index_t(1) << 32) && !fa_sort);
    if (by_count && !by_cand) frac_a_counted(s, false);
    // stable sort of the CSR entries by column alone (LSD radix), then one
    // thread per column run
    auto frac_a_sorted = [&](const double *area_src) {
        AXIS_HIP_CHECK(hipMemsetAsync(M->frac_a.get(), 0, sizeof(double) * std::max<index_t>(n_src, 1), s));
        if (nnz > 0) {
            DeviceBuffer<std::uint32_t> ckey(nz1);
            column_keys_kernel<<<grid_for(nnz, 256), 256, 0, s>>>(M->col.get(), nnz, ckey.get(), perm.get());
            AXIS_LAUNCH_CHECK();
            radix_sort_pairs_u32(ckey.get(), perm.get(), nnz, col_bits, s);
            frac_a_kernel<<<grid_for(nnz, 256), 256, 0, s>>>(ckey.get(), perm.get(), nnz, ov.get(), area_src, M->frac_a.get());
            AXIS_LAUNCH_CHECK();
        }
    };
    if (!by_count) frac_a_sorted(sg.d_area.get());
    if (by_cand) AXIS_HIP_CHECK(hipStreamWaitEvent(s, fev.done, 0));
    ev.rec(6, s);
    // the cell areas pass to the matrix (the geometry is done with them)
    M->area_a = std::move(sg.d_area);
    M->area_b = std::move(dg.d_area);
    if (M->area_a.size() < static_cast<std::size_t>(std::max<index_t>(n_src, 1))) M->area_a.resize(std::max<index_t>(n_src, 1));
    if (M->area_b.size() < static_cast<std::size_t>(std::max<index_t>(n_dst, 1))) M->area_b.resize(std::max<index_t>(n_dst, 1));
    PinnedArray<unsigned long long> h_tail;
    h_tail.resize(kH + 2);
    AXIS_HIP_CHECK(hipMemcpyAsync(h_tail.data(), max_row.get(), sizeof(int), hipMemcpyDeviceToHost, s));
    AXIS_HIP_CHECK(hipMemcpyAsync(h_tail.data() + kH + 1, col_flags.get(), 2 * sizeof(int), hipMemcpyDeviceToHost, s));
    if (grouped) AXIS_HIP_CHECK(hipMemcpyAsync(h_tail.data() + 1, d_hist.get(), sizeof(unsigned long long) * kH, hipMemcpyDeviceToHost, s));
    tr.mark("csr");
    if (cfg.unmapped == UnmappedAction::Error && n_dst > 0) {
        DeviceBuffer<unsigned long long> first(1);
        unsigned long long init = ~0ull, h_first = 0;
        AXIS_HIP_CHECK(hipMemcpyAsync(first.get(), &init, sizeof init, hipMemcpyHostToDevice, s));
        unmapped_kernel<<<grid_for(n_dst, 256), 256, 0, s>>>(M->row_ptr.get(), n_dst, dg.d_flags.get(), first.get());
        AXIS_LAUNCH_CHECK();
        AXIS_HIP_CHECK(hipMemcpyAsync(&h_first, first.get(), sizeof h_first, hipMemcpyDeviceToHost, s));
        AXIS_HIP_CHECK(hipStreamSynchronize(s));
        if (h_first != ~0ull)
            throw RuntimeError("WeightGenerator::generate_conservative: unmapped destination cell " + std::to_string(h_first));
    }
    AXIS_HIP_CHECK(hipStreamSynchronize(s));
    if (by_count && (h_tail[kH + 1] >> 32) != 0) {  // a column of more than kColLds entries
        frac_a_sorted(M->area_a.get());
        AXIS_HIP_CHECK(hipStreamSynchronize(s));
    }
    M->phase_ms[0] = ev.ms(0, 1);
    M->phase_ms[1] = ev.ms(2, 3);
    M->phase_ms[2] = ev.ms(3, 4);
    M->phase_ms[3] = ev.ms(5, 6);
    M->max_row_nnz = static_cast<int>(h_tail[0] & 0xffffffffu);
    if (grouped) {
        for (int l = 0; l < kH; ++l) M->len_hist[l] = static_cast<index_t>(h_tail[1 + l]);
        M->len_hist_ready = true;
    }
    M->gen_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return M;
}

}  // namespace axis_hip::detail
