// apply_kernels.hpp -- device code of the sparse apply, shared by the
// single-matrix launch (apply.hip) and the persistent multi-pair batch
// (batch.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <map>
#include <mutex>
#include <tuple>

#include "axis_hip.hpp"

namespace axis_hip::apply_detail {

constexpr int kLongRow = 64;  // rows with more nonzeros go wave-per-row
constexpr int kWave = 64;

/// Launch shape: T fields per lane, R rows per lane (rows-per-wavefront = 64 R).
struct Launch {
    int tile = 8;
    int rows_per_lane = 1;
};

__device__ inline float ld_f32(const float *p) { return *p; }
__device__ inline float ld_f32(const std::uint16_t *p) { return __uint_as_float(static_cast<unsigned>(*p) << 16); }

template <class YT>
__device__ inline void st(YT *p, double v) {
    *p = static_cast<YT>(v);
}

/// Compensated FP32 accumulator: s + c carries the sum; products are split by
/// TwoProduct (fmaf residual), the running sum by TwoSum.
struct Acc32 {
    float s = 0.0f, c = 0.0f;
    __device__ inline void add(float h, float l, float x) {
        float p = h * x;
        float pe = fmaf(h, x, -p);
        float t = s + p;
        float bp = t - s;
        float err = (s - (t - bp)) + (p - bp);
        s = t;
        c += fmaf(l, x, err + pe);
    }
    __device__ inline void merge(float s2, float c2) {
        float t = s + s2;
        float bp = t - s;
        float err = (s - (t - bp)) + (s2 - bp);
        s = t;
        c = c + (c2 + err);
    }
    __device__ inline double value() const { return static_cast<double>(s) + static_cast<double>(c); }
};

/// Matrix operands of one apply.
struct Mat {
    const index_t *row_ptr;
    const std::int32_t *col;
    const double *val;
    const float *hi, *lo;
    index_t n_src, n_dst;
};

/// Rows [r0, r0 + 64 R) of one wave (lane owns rows r0 + lane + 64 q), fields
/// [f0, f0 + nf), nf <= T.  Rows longer than kLongRow are skipped.
template <int T, bool MIXED, class XT, class YT>
__device__ inline void rows_tile(const Mat &A, const XT *X, YT *Y, index_t r0, int R, index_t f0, int nf, int lane) {
    for (int q = 0; q < R; ++q) {
        const index_t r = r0 + lane + static_cast<index_t>(q) * kWave;
        if (r >= A.n_dst) return;
        const index_t b = A.row_ptr[r], e = A.row_ptr[r + 1];
        if (e - b > kLongRow) continue;
        if constexpr (MIXED) {
            Acc32 acc[T];
            for (index_t idx = b; idx < e; ++idx) {
                const index_t c = A.col[idx];
                const float h = A.hi[idx], l = A.lo[idx];
#pragma unroll
                for (int t = 0; t < T; ++t)
                    if (t < nf) acc[t].add(h, l, ld_f32(X + c + (f0 + t) * A.n_src));
            }
#pragma unroll
            for (int t = 0; t < T; ++t)
                if (t < nf) st(Y + r + (f0 + t) * A.n_dst, acc[t].value());
        } else {
            double acc[T];
#pragma unroll
            for (int t = 0; t < T; ++t) acc[t] = 0.0;
            for (index_t idx = b; idx < e; ++idx) {
                const index_t c = A.col[idx];
                const double w = A.val[idx];
#pragma unroll
                for (int t = 0; t < T; ++t)
                    if (t < nf) acc[t] = fma(w, reinterpret_cast<const double *>(X)[c + (f0 + t) * A.n_src], acc[t]);
            }
#pragma unroll
            for (int t = 0; t < T; ++t)
                if (t < nf) reinterpret_cast<double *>(Y)[r + (f0 + t) * A.n_dst] = acc[t];
        }
    }
}

/// One long row by a whole wave: lanes stride over the row, partial sums are
/// combined by a butterfly of cross-lane shuffles (DPP / ds_swizzle).
template <int T, bool MIXED, class XT, class YT>
__device__ inline void long_row(const Mat &A, const XT *X, YT *Y, index_t r, index_t f0, int nf, int lane) {
    const index_t b = A.row_ptr[r], e = A.row_ptr[r + 1];
    if constexpr (MIXED) {
        Acc32 acc[T];
        for (index_t idx = b + lane; idx < e; idx += kWave) {
            const index_t c = A.col[idx];
            const float h = A.hi[idx], l = A.lo[idx];
#pragma unroll
            for (int t = 0; t < T; ++t)
                if (t < nf) acc[t].add(h, l, ld_f32(X + c + (f0 + t) * A.n_src));
        }
#pragma unroll
        for (int t = 0; t < T; ++t) {
            for (int off = kWave / 2; off >= 1; off >>= 1) {
                float s2 = __shfl_xor(acc[t].s, off, kWave);
                float c2 = __shfl_xor(acc[t].c, off, kWave);
                acc[t].merge(s2, c2);
            }
            if (lane == t && t < nf) st(Y + r + (f0 + t) * A.n_dst, acc[t].value());
        }
    } else {
        double acc[T];
#pragma unroll
        for (int t = 0; t < T; ++t) acc[t] = 0.0;
        for (index_t idx = b + lane; idx < e; idx += kWave) {
            const index_t c = A.col[idx];
            const double w = A.val[idx];
#pragma unroll
            for (int t = 0; t < T; ++t)
                if (t < nf) acc[t] = fma(w, reinterpret_cast<const double *>(X)[c + (f0 + t) * A.n_src], acc[t]);
        }
#pragma unroll
        for (int t = 0; t < T; ++t) {
            for (int off = kWave / 2; off >= 1; off >>= 1) acc[t] += __shfl_xor(acc[t], off, kWave);
            if (lane == t && t < nf) reinterpret_cast<double *>(Y)[r + (f0 + t) * A.n_dst] = acc[t];
        }
    }
}

/// Hardware places workgroup b on XCD b % 8; give XCD x a contiguous range of
/// work items so consecutive row blocks of one field tile share its L2.
__device__ inline unsigned xcd_remap(unsigned bid, unsigned n) {
    const unsigned xcd = bid & 7u, idx = bid >> 3;
    const unsigned q = n >> 3, rem = n & 7u;
    return xcd * q + (xcd < rem ? xcd : rem) + idx;
}

template <int T, bool MIXED, class XT, class YT>
__global__ void __launch_bounds__(kWave) spmm_rows_kernel(Mat A, const XT *X, YT *Y, index_t k, int R, unsigned n_row_blocks) {
    const unsigned w = xcd_remap(blockIdx.x, gridDim.x);
    const unsigned rb = w % n_row_blocks, tile = w / n_row_blocks;
    const index_t f0 = static_cast<index_t>(tile) * T;
    const int nf = static_cast<int>(k - f0 < T ? k - f0 : T);
    rows_tile<T, MIXED>(A, X, Y, static_cast<index_t>(rb) * kWave * R, R, f0, nf, threadIdx.x);
}

template <int T, bool MIXED, class XT, class YT>
__global__ void __launch_bounds__(kWave) spmm_long_kernel(Mat A, const XT *X, YT *Y, index_t k, const std::int32_t *rows, int n_rows) {
    const unsigned w = blockIdx.x;
    const int li = static_cast<int>(w % static_cast<unsigned>(n_rows));
    const index_t f0 = static_cast<index_t>(w / static_cast<unsigned>(n_rows)) * T;
    const int nf = static_cast<int>(k - f0 < T ? k - f0 : T);
    long_row<T, MIXED>(A, X, Y, rows[li], f0, nf, threadIdx.x);
}

inline Mat mat_of(const InterpolationMatrix &A) {
    return Mat{A.row_ptr.get(), A.col32.get(), A.val.get(), A.val_hi.get(), A.val_lo.get(), A.n_src, A.n_dst};
}

template <int T, bool MIXED, class XT, class YT>
void launch_t(const InterpolationMatrix &A, const void *x, void *y, index_t k, int R, hipStream_t s) {
    const Mat M = mat_of(A);
    const index_t rows_per_block = static_cast<index_t>(kWave) * R;
    const unsigned nrb = static_cast<unsigned>((A.n_dst + rows_per_block - 1) / rows_per_block);
    const unsigned ntiles = static_cast<unsigned>((k + T - 1) / T);
    const std::uint64_t nwork = static_cast<std::uint64_t>(nrb) * ntiles;
    if (nwork >= (1ull << 31)) throw InvalidArgument("apply: launch too large");
    spmm_rows_kernel<T, MIXED, XT, YT><<<static_cast<unsigned>(nwork), kWave, 0, s>>>(M, static_cast<const XT *>(x), static_cast<YT *>(y), k, R, nrb);
    AXIS_LAUNCH_CHECK();
    if (A.n_long_rows > 0) {
        const std::uint64_t nl = static_cast<std::uint64_t>(A.n_long_rows) * ntiles;
        spmm_long_kernel<T, MIXED, XT, YT><<<static_cast<unsigned>(nl), kWave, 0, s>>>(M, static_cast<const XT *>(x), static_cast<YT *>(y), k,
                                                                                       A.long_rows.get(), A.n_long_rows);
        AXIS_LAUNCH_CHECK();
    }
}

template <bool MIXED, class XT, class YT>
void launch_xy(const InterpolationMatrix &A, const void *x, void *y, index_t k, const Launch &L, hipStream_t s) {
    switch (L.tile) {
        case 1: return launch_t<1, MIXED, XT, YT>(A, x, y, k, L.rows_per_lane, s);
        case 2: return launch_t<2, MIXED, XT, YT>(A, x, y, k, L.rows_per_lane, s);
        case 4: return launch_t<4, MIXED, XT, YT>(A, x, y, k, L.rows_per_lane, s);
        case 8: return launch_t<8, MIXED, XT, YT>(A, x, y, k, L.rows_per_lane, s);
        case 16: return launch_t<16, MIXED, XT, YT>(A, x, y, k, L.rows_per_lane, s);
        default: throw InvalidArgument("apply: unsupported field tile");
    }
}

inline void launch(const InterpolationMatrix &A, const void *x, FieldType xt, void *y, FieldType yt, index_t k, ApplyMode mode, const Launch &L,
                   hipStream_t s) {
    if (mode == ApplyMode::Plain) return launch_xy<false, double, double>(A, x, y, k, L, s);
    if (xt == FieldType::F32 && yt == FieldType::F32) return launch_xy<true, float, float>(A, x, y, k, L, s);
    if (xt == FieldType::F32 && yt == FieldType::F64) return launch_xy<true, float, double>(A, x, y, k, L, s);
    if (xt == FieldType::BF16 && yt == FieldType::F32) return launch_xy<true, std::uint16_t, float>(A, x, y, k, L, s);
    if (xt == FieldType::BF16 && yt == FieldType::F64) return launch_xy<true, std::uint16_t, double>(A, x, y, k, L, s);
    throw InvalidArgument("apply: unsupported field type combination");
}

// ------------------------------------------------------------- autotuner
using TuneKey = std::tuple<index_t, index_t, index_t, int, int, int>;  // n_dst, nnz, k, xt, yt, mode

inline std::map<TuneKey, Launch> &tune_cache() {
    static std::map<TuneKey, Launch> m;
    return m;
}
inline std::mutex &tune_mutex() {
    static std::mutex m;
    return m;
}

inline Launch heuristic(index_t k) {
    Launch L;
    L.tile = k >= 16 ? 16 : (k >= 8 ? 8 : (k >= 4 ? 4 : (k >= 2 ? 2 : 1)));
    L.rows_per_lane = 1;
    return L;
}

inline Launch choose(const InterpolationMatrix &A, index_t k, FieldType xt, FieldType yt, ApplyMode mode) {
    std::lock_guard<std::mutex> g(tune_mutex());
    auto it = tune_cache().find(TuneKey{A.n_dst, A.nnz, k, static_cast<int>(xt), static_cast<int>(yt), static_cast<int>(mode)});
    if (it != tune_cache().end()) return it->second;
    return heuristic(k);
}

}  // namespace axis_hip::apply_detail
