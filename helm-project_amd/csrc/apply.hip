// apply.hip -- sparse weight apply Y = A X on layout_left fields.
//
// Replaces apply / batch_apply (libs/axis/include/axis/solver/apply.hpp:61-119,
// 275-343), whose CSR branch calls KokkosSparse::spmv.  Two variants:
//
//   plain  FP64 fields, FP64 weights, FP64 accumulation (the plain GPU method);
//   mixed  FP32 or BF16 fields, weights split into FP32 hi + lo, FP32
//          TwoProduct/TwoSum-compensated accumulation, FP32 or FP64 output.
//
// The kernels live in apply_kernels.hpp: rows cut into segments of at most E
// nonzeros, one lane per segment with its weights held in registers while the
// fields stream through, long rows finished by a deterministic reduction of
// their segment partials.  This file holds the per-matrix set-up (FP32 hi/lo
// split, int32 columns, segments), the entry points, the autotuner over
// (fields per step, field chunks), the conservation integrals and the COO
// constructor.
#include <hip/hip_bf16.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>

#include "apply_kernels.hpp"
#include "axis_hip.hpp"
#include "host_util.hpp"

namespace axis_hip {

namespace {

__global__ void split_kernel(const double *v, const index_t *col, index_t nnz, float *hi, float *lo, std::int32_t *c32) {
    index_t e = static_cast<index_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (e >= nnz) return;
    double x = v[e];
    float h = static_cast<float>(x);
    hi[e] = h;
    lo[e] = static_cast<float>(x - static_cast<double>(h));
    c32[e] = static_cast<std::int32_t>(col[e]);
}

}  // namespace

void InterpolationMatrix::prepare_apply(hipStream_t s) {
    if (col32.size() >= static_cast<std::size_t>(std::max<index_t>(nnz, 1)) && val_hi.size() > 0) return;
    if (n_src >= (index_t(1) << 31)) throw InvalidArgument("apply: more than 2^31 source cells");
    const index_t n1 = std::max<index_t>(nnz, 1);
    col32.resize(n1);
    val_hi.resize(n1);
    val_lo.resize(n1);
    if (nnz > 0) {
        split_kernel<<<grid_for(nnz, 256), 256, 0, s>>>(val.get(), col.get(), nnz, val_hi.get(), val_lo.get(), col32.get());
        AXIS_LAUNCH_CHECK();
    }
    // E = the smallest register segment (4, 8, 12, 16) that leaves at most
    // 0.1% of the rows longer than one segment
    std::vector<index_t> rp(n_dst + 1);
    row_ptr.download(rp.data(), n_dst + 1, s);
    AXIS_HIP_CHECK(hipStreamSynchronize(s));
    index_t hist[apply_detail::kMaxE + 2] = {0};
    for (index_t r = 0; r < n_dst; ++r) {
        index_t len = rp[r + 1] - rp[r];
        hist[len > apply_detail::kMaxE ? apply_detail::kMaxE + 1 : len]++;
    }
    seg_cap = apply_detail::kMaxE;
    for (int cap : {4, 8, 12, 16}) {
        index_t above = 0;
        for (int l = cap + 1; l <= apply_detail::kMaxE + 1; ++l) above += hist[l];
        if (above * 1000 <= n_dst) {
            seg_cap = cap;
            break;
        }
    }
    // rows are processed in index order: a destination mesh numbered along
    // rows of cells (cubed-sphere faces, lat-lon rows) gives each wave
    // neighbouring cells, whose source columns share cache lines, and full-line
    // stores of Y; regrouping rows by source column was measured slower
    std::vector<std::int32_t> order(n_dst);
    for (index_t r = 0; r < n_dst; ++r) order[r] = static_cast<std::int32_t>(r);
    std::vector<std::int32_t> h_row, h_len, h_slot, h_lrow, h_lslot;
    std::vector<index_t> h_begin;
    h_row.reserve(n_dst);
    h_len.reserve(n_dst);
    h_slot.reserve(n_dst);
    h_begin.reserve(n_dst);
    int slots = 0;
    for (std::int32_t r : order) {
        const index_t b = rp[r], e = rp[r + 1];
        if (e == b) continue;  // empty rows are appended after the others
        if (e - b <= seg_cap) {
            h_row.push_back(r);
            h_begin.push_back(b);
            h_len.push_back(static_cast<std::int32_t>(e - b));
            h_slot.push_back(-1);
        } else {
            h_lrow.push_back(r);
            h_lslot.push_back(slots);
            for (index_t q = b; q < e; q += seg_cap) {
                h_row.push_back(r);
                h_begin.push_back(q);
                h_len.push_back(static_cast<std::int32_t>(std::min<index_t>(seg_cap, e - q)));
                h_slot.push_back(slots++);
            }
        }
    }
    // empty rows: a zero-length segment writes y = 0 (the reference's spmv
    // overwrites every destination entry)
    for (std::int32_t r : order)
        if (rp[r + 1] == rp[r]) {
            h_row.push_back(r);
            h_begin.push_back(rp[r]);
            h_len.push_back(0);
            h_slot.push_back(-1);
        }
    h_lslot.push_back(slots);
    n_seg = static_cast<index_t>(h_row.size());
    n_long_rows = static_cast<int>(h_lrow.size());
    n_slots = slots;
    auto put32 = [&](DeviceBuffer<std::int32_t> &d, const std::vector<std::int32_t> &h) {
        d.resize(std::max<std::size_t>(h.size(), 1));
        if (!h.empty()) d.upload(h.data(), h.size(), s);
    };
    put32(seg_row, h_row);
    put32(seg_len, h_len);
    put32(seg_slot, h_slot);
    put32(long_rows, h_lrow);
    put32(long_slot, h_lslot);
    seg_begin.resize(std::max<std::size_t>(h_begin.size(), 1));
    if (!h_begin.empty()) seg_begin.upload(h_begin.data(), h_begin.size(), s);
    AXIS_HIP_CHECK(hipStreamSynchronize(s));
}

void batch_apply(InterpolationMatrix &A, const void *x, FieldType xt, void *y, FieldType yt, index_t n_src, index_t n_dst, index_t k, ApplyMode mode,
                 hipStream_t s) {
    // apply.hpp:63-71 / :285-296: validate before touching dst
    if (n_src != A.n_src) throw InvalidArgument("axis::solver::batch_apply: src.extent(0) (" + std::to_string(n_src) + ") != matrix.n_src() (" +
                                                std::to_string(A.n_src) + ")");
    if (n_dst != A.n_dst) throw InvalidArgument("axis::solver::batch_apply: dst.extent(0) (" + std::to_string(n_dst) + ") != matrix.n_dst() (" +
                                                std::to_string(A.n_dst) + ")");
    if (k < 0) throw InvalidArgument("batch_apply: negative field count");
    if (mode == ApplyMode::Plain && (xt != FieldType::F64 || yt != FieldType::F64))
        throw InvalidArgument("batch_apply: the plain apply takes FP64 fields");
    if (mode == ApplyMode::Mixed && (xt == FieldType::F64 || yt == FieldType::BF16))
        throw InvalidArgument("batch_apply: the mixed apply takes FP32/BF16 fields and writes FP32/FP64");
    if (k == 0 || n_dst == 0) return;
    A.prepare_apply(s);
    apply_detail::Launch L = apply_detail::choose(A, k, xt, yt, mode);
    apply_detail::launch(A, x, xt, y, yt, k, mode, L, s);
}

std::pair<int, int> autotune_apply(InterpolationMatrix &A, const void *x, FieldType xt, void *y, FieldType yt, index_t k, ApplyMode mode,
                                   hipStream_t s) {
    if (k <= 0) throw InvalidArgument("autotune_apply: no fields");
    A.prepare_apply(s);
    hipEvent_t e0, e1;
    AXIS_HIP_CHECK(hipEventCreate(&e0));
    AXIS_HIP_CHECK(hipEventCreate(&e1));
    apply_detail::Launch best = apply_detail::heuristic(A, k, mode);
    float best_ms = 1e30f;
    for (int fu : {1, 2, 4}) {
        if (fu > 1 && fu / 2 >= k) continue;
        for (int chunks : {1, 2, 4, 8, 16, 32, 64}) {
            if (chunks > 1 && (chunks / 2) * fu >= k) continue;
            apply_detail::Launch L;
            L.fu = fu;
            L.chunks = chunks;
            apply_detail::launch(A, x, xt, y, yt, k, mode, L, s);  // warm
            AXIS_HIP_CHECK(hipEventRecord(e0, s));
            const int reps = 3;
            for (int i = 0; i < reps; ++i) apply_detail::launch(A, x, xt, y, yt, k, mode, L, s);
            AXIS_HIP_CHECK(hipEventRecord(e1, s));
            AXIS_HIP_CHECK(hipEventSynchronize(e1));
            float ms = 0.0f;
            AXIS_HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
            if (ms < best_ms) {
                best_ms = ms;
                best = L;
            }
        }
    }
    AXIS_HIP_CHECK(hipEventDestroy(e0));
    AXIS_HIP_CHECK(hipEventDestroy(e1));
    {
        std::lock_guard<std::mutex> g(apply_detail::tune_mutex());
        apply_detail::tune_cache()[apply_detail::TuneKey{A.n_dst, A.nnz, k, static_cast<int>(xt), static_cast<int>(yt), static_cast<int>(mode)}] =
            best;
    }
    // leave the output of the chosen shape in y
    apply_detail::launch(A, x, xt, y, yt, k, mode, best, s);
    AXIS_HIP_CHECK(hipStreamSynchronize(s));
    return {best.fu, best.chunks};
}

// ------------------------------------------------------ conservation check
namespace {

constexpr int kRedBlock = 256;

// Neumaier-compensated FP64 partial sums of a[i] * b[i] * c[i] (c optional,
// terms skipped where gate[i] == 0 when a gate is given), one (sum, comp)
// pair per block.
__global__ void __launch_bounds__(kRedBlock) triple_sum_kernel(const double *a, const double *b, const double *c, const double *gate, index_t n,
                                                               double *partial) {
    __shared__ double ls[kRedBlock], lc[kRedBlock];
    double s = 0.0, comp = 0.0;
    for (index_t i = static_cast<index_t>(blockIdx.x) * kRedBlock + threadIdx.x; i < n; i += static_cast<index_t>(gridDim.x) * kRedBlock) {
        if (gate && gate[i] == 0.0) continue;
        double t = a[i] * b[i];
        if (c) t *= c[i];
        double u = s + t;
        comp += (fabs(s) >= fabs(t)) ? (s - u) + t : (t - u) + s;
        s = u;
    }
    ls[threadIdx.x] = s;
    lc[threadIdx.x] = comp;
    __syncthreads();
    for (int h = kRedBlock / 2; h > 0; h >>= 1) {
        if (threadIdx.x < h) {
            double x = ls[threadIdx.x], y = ls[threadIdx.x + h];
            double u = x + y;
            double e = (fabs(x) >= fabs(y)) ? (x - u) + y : (y - u) + x;
            ls[threadIdx.x] = u;
            lc[threadIdx.x] += lc[threadIdx.x + h] + e;
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        partial[2 * blockIdx.x] = ls[0];
        partial[2 * blockIdx.x + 1] = lc[0];
    }
}

double compensated_triple_sum(const double *a, const double *b, const double *c, const double *gate, index_t n, hipStream_t s) {
    if (n <= 0) return 0.0;
    const unsigned nb = static_cast<unsigned>(std::min<index_t>((n + kRedBlock - 1) / kRedBlock, 1024));
    DeviceBuffer<double> part(2 * nb);
    triple_sum_kernel<<<nb, kRedBlock, 0, s>>>(a, b, c, gate, n, part.get());
    AXIS_LAUNCH_CHECK();
    std::vector<double> h(2 * nb);
    part.download(h.data(), 2 * nb, s);
    AXIS_HIP_CHECK(hipStreamSynchronize(s));
    double sum = 0.0, comp = 0.0;
    for (unsigned i = 0; i < nb; ++i) {
        double t = h[2 * i];
        double u = sum + t;
        comp += (std::fabs(sum) >= std::fabs(t)) ? (sum - u) + t : (t - u) + sum;
        sum = u;
        comp += h[2 * i + 1];
    }
    return sum + comp;
}

}  // namespace

// check_conservation, conservation.cpp:23-103
ConservationReport check_conservation(const InterpolationMatrix &A, const double *d_src, const double *d_dst, NormType norm, hipStream_t s) {
    ConservationReport r;
    r.src_integral = compensated_triple_sum(d_src, A.area_a.get(), A.frac_a.get(), nullptr, A.n_src, s);
    if (norm == NormType::DstArea)
        r.dst_integral = compensated_triple_sum(d_dst, A.area_b.get(), nullptr, A.frac_b.get(), A.n_dst, s);
    else
        r.dst_integral = compensated_triple_sum(d_dst, A.area_b.get(), A.frac_b.get(), nullptr, A.n_dst, s);
    r.absolute_error = std::fabs(r.src_integral - r.dst_integral);
    const double den = std::max({std::fabs(r.src_integral), std::fabs(r.dst_integral), std::numeric_limits<double>::epsilon()});
    r.relative_error = r.absolute_error / den;
    return r;
}

// ------------------------------------------------------ COO constructor
std::unique_ptr<InterpolationMatrix> matrix_from_coo(const index_t *rows, const index_t *cols, const double *vals, index_t nnz, index_t n_src,
                                                     index_t n_dst, const double *frac_a, const double *frac_b, const double *area_a,
                                                     const double *area_b) {
    if (nnz < 0 || n_src < 0 || n_dst < 0) throw InvalidArgument("matrix_from_coo: negative size");
    for (index_t k = 0; k < nnz; ++k) {
        if (rows[k] < 0 || rows[k] >= n_dst) throw InvalidArgument("matrix_from_coo: row index out of range");
        if (cols[k] < 0 || cols[k] >= n_src) throw InvalidArgument("matrix_from_coo: column index out of range");
    }
    // to_csr (interpolation_matrix.hpp:180-255): (row, col) order, duplicates kept
    std::vector<index_t> perm(nnz);
    for (index_t k = 0; k < nnz; ++k) perm[k] = k;
    std::stable_sort(perm.begin(), perm.end(), [&](index_t a, index_t b) {
        if (rows[a] != rows[b]) return rows[a] < rows[b];
        return cols[a] < cols[b];
    });
    std::vector<index_t> rp(n_dst + 1, 0), c(nnz);
    std::vector<double> v(nnz);
    for (index_t k = 0; k < nnz; ++k) rp[rows[perm[k]] + 1]++;
    for (index_t i = 0; i < n_dst; ++i) rp[i + 1] += rp[i];
    int max_row = 0;
    for (index_t i = 0; i < n_dst; ++i) max_row = std::max<int>(max_row, static_cast<int>(rp[i + 1] - rp[i]));
    for (index_t k = 0; k < nnz; ++k) {
        c[k] = cols[perm[k]];
        v[k] = vals[perm[k]];
    }
    auto M = std::make_unique<InterpolationMatrix>();
    M->n_src = n_src;
    M->n_dst = n_dst;
    M->nnz = nnz;
    M->max_row_nnz = max_row;
    hipStream_t s = nullptr;
    M->row_ptr.resize(n_dst + 1);
    M->row_ptr.upload(rp.data(), n_dst + 1, s);
    M->col.resize(std::max<index_t>(nnz, 1));
    M->col.upload(c.data(), nnz, s);
    M->val.resize(std::max<index_t>(nnz, 1));
    M->val.upload(v.data(), nnz, s);
    auto put = [&](DeviceBuffer<double> &d, const double *h, index_t n, double fill) {
        d.resize(std::max<index_t>(n, 1));
        std::vector<double> tmp(n, fill);
        if (h) std::copy(h, h + n, tmp.begin());
        d.upload(tmp.data(), n, s);
    };
    put(M->frac_a, frac_a, n_src, 1.0);
    put(M->frac_b, frac_b, n_dst, 1.0);
    put(M->area_a, area_a, n_src, 1.0);
    put(M->area_b, area_b, n_dst, 1.0);
    AXIS_HIP_CHECK(hipStreamSynchronize(s));
    return M;
}

}  // namespace axis_hip
