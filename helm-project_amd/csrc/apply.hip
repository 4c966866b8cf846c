// apply.hip -- sparse weight apply Y = A X on layout_left fields.
//
// Replaces apply / batch_apply (libs/axis/include/axis/solver/apply.hpp:61-119,
// 275-343), whose CSR branch calls KokkosSparse::spmv.  Two variants:
//
//   plain  FP64 fields, FP64 weights, FP64 accumulation (the plain GPU method);
//   mixed  FP32 or BF16 fields, weights split into FP32 hi + lo, FP32
//          TwoProduct/TwoSum-compensated accumulation, FP32 or FP64 output.
//
// Mapping: one lane per destination row, a tile of T fields per lane (the row's
// (col, w) entries are loaded once and reused across the T fields), so each
// field's output column is written by 64 consecutive rows per wave (full-line
// stores) and the source gathers of neighbouring rows share cache lines.  Rows
// longer than kLongRow go to a wave-per-row kernel whose lanes split the row and
// combine partial sums with cross-lane (DPP/swizzle) reductions.  Work items are
// (row block, field tile) pairs remapped so that each XCD walks a contiguous
// range, keeping its source-column working set in its own L2.
#include <hip/hip_bf16.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>

#include "apply_kernels.hpp"
#include "axis_hip.hpp"

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

__global__ void long_rows_kernel(const index_t *row_ptr, index_t n_dst, int limit, std::int32_t *list, int *n_list) {
    index_t r = static_cast<index_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (r >= n_dst) return;
    if (row_ptr[r + 1] - row_ptr[r] > limit) list[atomicAdd(n_list, 1)] = static_cast<std::int32_t>(r);
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
    // rows handled by the wave-per-row kernel
    DeviceBuffer<std::int32_t> tmp(std::max<index_t>(n_dst, 1));
    DeviceBuffer<int> cnt(1);
    AXIS_HIP_CHECK(hipMemsetAsync(cnt.get(), 0, sizeof(int), s));
    if (n_dst > 0) {
        long_rows_kernel<<<grid_for(n_dst, 256), 256, 0, s>>>(row_ptr.get(), n_dst, apply_detail::kLongRow, tmp.get(), cnt.get());
        AXIS_LAUNCH_CHECK();
    }
    int h = 0;
    AXIS_HIP_CHECK(hipMemcpyAsync(&h, cnt.get(), sizeof(int), hipMemcpyDeviceToHost, s));
    AXIS_HIP_CHECK(hipStreamSynchronize(s));
    n_long_rows = h;
    long_rows.resize(std::max(h, 1));
    if (h > 0) {
        std::vector<std::int32_t> v(h);
        tmp.download(v.data(), h, s);
        AXIS_HIP_CHECK(hipStreamSynchronize(s));
        std::sort(v.begin(), v.end());
        long_rows.upload(v.data(), h, s);
        AXIS_HIP_CHECK(hipStreamSynchronize(s));
    }
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
    A.prepare_apply(s);
    hipEvent_t e0, e1;
    AXIS_HIP_CHECK(hipEventCreate(&e0));
    AXIS_HIP_CHECK(hipEventCreate(&e1));
    apply_detail::Launch best = apply_detail::heuristic(k);
    float best_ms = 1e30f;
    const int tiles[] = {1, 2, 4, 8, 16};
    const int rpl[] = {1, 2, 4};
    for (int t : tiles) {
        if (t > 1 && t / 2 >= k) continue;  // a tile no smaller one fills as well
        for (int r : rpl) {
            apply_detail::Launch L;
            L.tile = t;
            L.rows_per_lane = r;
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
    return {best.tile, best.rows_per_lane};
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
