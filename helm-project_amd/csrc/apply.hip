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
#include <map>
#include <mutex>

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

// histogram of row lengths (0 .. kMaxE, and "longer"): per-block LDS counts,
// then one global atomic per bin and block
__global__ void len_hist_kernel(const index_t *rp, index_t n_dst, index_t *hist) {
    constexpr int kH = apply_detail::kMaxE + 2;
    __shared__ unsigned cnt[kH];
    if (threadIdx.x < kH) cnt[threadIdx.x] = 0;
    __syncthreads();
    index_t r = static_cast<index_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (r < n_dst) {
        const index_t len = rp[r + 1] - rp[r];
        atomicAdd(&cnt[len > apply_detail::kMaxE ? apply_detail::kMaxE + 1 : len], 1u);
    }
    __syncthreads();
    if (threadIdx.x < kH && cnt[threadIdx.x])
        atomicAdd(reinterpret_cast<unsigned long long *>(hist + threadIdx.x), static_cast<unsigned long long>(cnt[threadIdx.x]));
}

// sort key of row r: (coarse spatial cell << 32) | r, the cell taken from
// the first row of r's aligned block of 32 rows, so that blocks stay whole
// and Y is written in full 128-byte lines (FP32); rows without a key keep
// index order
__global__ void order_key_kernel(const std::uint32_t *row_key, int shift, index_t n_dst, std::uint64_t *key) {
    index_t r = static_cast<index_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (r >= n_dst) return;
    const std::uint64_t coarse = row_key ? static_cast<std::uint64_t>(row_key[r & ~index_t(31)] >> shift) : 0ull;
    key[r] = (coarse << 32) | static_cast<std::uint64_t>(r);
}

__global__ void seg_count_kernel(const std::uint64_t *key, const index_t *rp, index_t n_dst, int E, index_t *n_seg, index_t *n_slot,
                                 index_t *n_long) {
    index_t p = static_cast<index_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (p >= n_dst) return;
    const index_t r = static_cast<index_t>(key[p] & 0xFFFFFFFFull);
    const index_t len = rp[r + 1] - rp[r];
    const bool lng = len > E;
    const index_t ns = lng ? (len + E - 1) / E : 1;
    n_seg[p] = ns;
    n_slot[p] = lng ? ns : 0;
    n_long[p] = lng ? 1 : 0;
}

__global__ void seg_fill_kernel(const std::uint64_t *key, const index_t *rp, index_t n_dst, int E, const index_t *seg_off, const index_t *slot_off,
                                const index_t *long_off, std::int32_t *seg_row, index_t *seg_begin, std::int32_t *seg_len, std::int32_t *seg_slot,
                                std::int32_t *long_rows, std::int32_t *long_slot) {
    index_t p = static_cast<index_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (p >= n_dst) return;
    const index_t r = static_cast<index_t>(key[p] & 0xFFFFFFFFull);
    const index_t b = rp[r], len = rp[r + 1] - b;
    const bool lng = len > E;
    const index_t s0 = seg_off[p];
    if (!lng) {
        seg_row[s0] = static_cast<std::int32_t>(r);
        seg_begin[s0] = b;
        seg_len[s0] = static_cast<std::int32_t>(len);
        seg_slot[s0] = -1;
        return;
    }
    const index_t ns = (len + E - 1) / E;
    for (index_t q = 0; q < ns; ++q) {
        seg_row[s0 + q] = static_cast<std::int32_t>(r);
        seg_begin[s0 + q] = b + q * E;
        seg_len[s0 + q] = static_cast<std::int32_t>(len - q * E < E ? len - q * E : E);
        seg_slot[s0 + q] = static_cast<std::int32_t>(slot_off[p] + q);
    }
    long_rows[long_off[p]] = static_cast<std::int32_t>(r);
    long_slot[long_off[p]] = static_cast<std::int32_t>(slot_off[p]);
}

// Direct plain apply of a matrix that has no apply layout yet (a few FP64
// fields): the same sums as the segmented apply, without building it.  A row
// of at most E nonzeros is one FMA chain in CSR order (one thread); a longer
// row is cut into E-entry segments, each an FMA chain from zero, and the
// segment sums are added in order from zero (one wave per row), exactly as
// the segmented kernels and long_reduce_kernel do.
// (E a template constant: a row's column and weight loads, then its field
// gathers, are all issued before the FMA chain needs them)
template <int E>
__global__ void __launch_bounds__(256) direct_rows_kernel(const index_t *__restrict__ rp, const index_t *__restrict__ col,
                                                          const double *__restrict__ val, index_t n_dst, index_t n_src, const double *__restrict__ X,
                                                          double *__restrict__ Y, index_t k, std::int32_t *__restrict__ long_list,
                                                          int *__restrict__ n_long) {
    const index_t r = static_cast<index_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (r >= n_dst) return;
    const index_t b = rp[r];
    const int len = static_cast<int>(min<index_t>(rp[r + 1] - b, E + 1));
    if (len > E) {
        long_list[atomicAdd(n_long, 1)] = static_cast<std::int32_t>(r);
        return;
    }
    index_t c[E];
    double v[E];
#pragma unroll
    for (int q = 0; q < E; ++q) {
        c[q] = 0;
        v[q] = 0.0;
        if (q < len) {
            c[q] = col[b + q];
            v[q] = val[b + q];
        }
    }
    for (index_t f = 0; f < k; ++f) {
        const double *xf = X + f * n_src;
        double xv[E];
#pragma unroll
        for (int q = 0; q < E; ++q) xv[q] = q < len ? xf[c[q]] : 0.0;
        double acc = 0.0;
#pragma unroll
        for (int q = 0; q < E; ++q)
            if (q < len) acc = fma(v[q], xv[q], acc);
        Y[f * n_dst + r] = acc;
    }
}

__global__ void __launch_bounds__(256) direct_long_kernel(const index_t *__restrict__ rp, const index_t *__restrict__ col,
                                                          const double *__restrict__ val, index_t n_dst, index_t n_src, const double *__restrict__ X,
                                                          double *__restrict__ Y, index_t k, int E, const std::int32_t *__restrict__ long_list,
                                                          int n_long) {
    const int lane = static_cast<int>(threadIdx.x) & 63;
    const index_t i = static_cast<index_t>(blockIdx.x) * 4 + threadIdx.x / 64;
    if (i >= n_long) return;  // whole wave (no workgroup barriers below)
    const index_t r = long_list[i];
    const index_t b = rp[r], len = rp[r + 1] - b;
    const index_t ns = (len + E - 1) / E;
    for (index_t f = 0; f < k; ++f) {
        const double *xf = X + f * n_src;
        double tot = 0.0;
        for (index_t q0 = 0; q0 < ns; q0 += 64) {
            const index_t q = q0 + lane;
            double part = 0.0;
            if (q < ns) {
                const index_t sb = b + q * E, se = min(sb + E, b + len);
                for (index_t e = sb; e < se; ++e) part = fma(val[e], xf[col[e]], part);
            }
            const int m = static_cast<int>(min<index_t>(64, ns - q0));
            for (int j = 0; j < m; ++j) tot += __shfl(part, j, 64);
        }
        if (lane == 0) Y[f * n_dst + r] = tot;
    }
}

// exclusive scan of one int per thread over a tile workgroup
__device__ inline int block_exclusive_scan(int v, int *wsum, int &total) {
    const int lane = static_cast<int>(threadIdx.x) & 63, wave = static_cast<int>(threadIdx.x) >> 6;
    int x = v;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        int y = __shfl_up(x, d, 64);
        if (lane >= d) x += y;
    }
    if (lane == 63) wsum[wave] = x;
    __syncthreads();
    int before = 0;
    total = 0;
    for (int w = 0; w < apply_detail::kTileSeg / 64; ++w) {
        if (w < wave) before += wsum[w];
        total += wsum[w];
    }
    __syncthreads();
    return before + x - v;
}

constexpr int pow2_at_least(int v) {
    int p = 1;
    while (p < v) p <<= 1;
    return p;
}

// Tile t = segments [256 t, 256 t + 256): sorts the aligned 4-column chunks
// its nonzeros touch in LDS (bitonic), keeps the distinct ones, and records
// for each nonzero the position of its column in the staged values
// (4 * chunk position + column % 4).
template <int E>
__global__ void __launch_bounds__(apply_detail::kTileSeg) tile_build_kernel(const index_t *seg_begin, const std::int32_t *seg_len, index_t n_seg,
                                                                          const std::int32_t *col32, std::int32_t *tile_nu, std::int32_t *tile_col,
                                                                          std::uint16_t *lcol) {
    constexpr int kB = apply_detail::kTileSeg;
    constexpr int SZ = pow2_at_least(kB * E);
    constexpr int C = SZ / kB;
    constexpr int kEmpty = 0x7FFFFFFF;
    __shared__ int key[SZ];
    __shared__ int uniq[apply_detail::kTileChunks];
    __shared__ int wsum[kB / 64];
    const index_t t = blockIdx.x;
    const int tid = static_cast<int>(threadIdx.x);
    const index_t i = t * kB + tid;
    const int n = i < n_seg ? seg_len[i] : 0;
    const index_t b = i < n_seg ? seg_begin[i] : 0;
    for (int q = tid; q < SZ; q += kB) key[q] = kEmpty;
    __syncthreads();
    for (int e = 0; e < n; ++e) key[tid * E + e] = col32[b + e] / apply_detail::kChunk;
    __syncthreads();
    for (int kk = 2; kk <= SZ; kk <<= 1) {
        for (int j = kk >> 1; j > 0; j >>= 1) {
            for (int q = tid; q < SZ; q += kB) {
                const int p = q ^ j;
                if (p > q) {
                    const int x = key[q], y = key[p];
                    const bool up = (q & kk) == 0;
                    if ((x > y) == up) {
                        key[q] = y;
                        key[p] = x;
                    }
                }
            }
            __syncthreads();
        }
    }
    int cnt = 0;
    for (int q = tid * C; q < tid * C + C; ++q) {
        const int v = key[q];
        if (v != kEmpty && (q == 0 || key[q - 1] != v)) ++cnt;
    }
    int total = 0;
    int pos = block_exclusive_scan(cnt, wsum, total);
    if (total > apply_detail::kTileChunks) {
        if (tid == 0) tile_nu[t] = -1;
        return;  // whole workgroup
    }
    for (int q = tid * C; q < tid * C + C; ++q) {
        const int v = key[q];
        if (v != kEmpty && (q == 0 || key[q - 1] != v)) {
            uniq[pos] = v;
            tile_col[t * apply_detail::kTileChunks + pos] = v * apply_detail::kChunk;
            ++pos;
        }
    }
    __syncthreads();
    for (int e = 0; e < n; ++e) {
        const int col = col32[b + e];
        const int c = col / apply_detail::kChunk;
        int lo = 0, hi = total;
        while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            if (uniq[mid] < c)
                lo = mid + 1;
            else
                hi = mid;
        }
        lcol[b + e] = static_cast<std::uint16_t>(lo * apply_detail::kChunk + col % apply_detail::kChunk);
    }
    if (tid == 0) tile_nu[t] = total;
}

// Wave tile t = segments [64 t, 64 t + 64): one wave sorts the aligned
// chunks (CW columns each) its nonzeros touch, keeps the distinct ones (at
// most kWaveChunks, else wnu[t] = -1) and records each nonzero's element
// position in the staged chunks (CW * chunk position + column % CW).
template <int E, int CW>
__global__ void __launch_bounds__(apply_detail::kWaveSeg) wave_tile_build_kernel(const index_t *seg_begin, const std::int32_t *seg_len, index_t n_seg,
                                                                                 const std::int32_t *col32, std::int32_t *wnu, std::int32_t *wcol,
                                                                                 std::uint16_t *lpos) {
    constexpr int kB = apply_detail::kWaveSeg;
    constexpr int SZ = pow2_at_least(kB * E);
    constexpr int C = SZ / kB;
    constexpr int kEmpty = 0x7FFFFFFF;
    __shared__ int key[SZ];
    __shared__ int uniq[apply_detail::kWaveChunks];
    const index_t t = blockIdx.x;
    const int tid = static_cast<int>(threadIdx.x);
    const index_t i = t * kB + tid;
    const int n = i < n_seg ? seg_len[i] : 0;
    const index_t b = i < n_seg ? seg_begin[i] : 0;
    for (int q = tid; q < SZ; q += kB) key[q] = kEmpty;
    __syncthreads();
    for (int e = 0; e < n; ++e) key[tid * E + e] = col32[b + e] / CW;
    __syncthreads();
    for (int kk = 2; kk <= SZ; kk <<= 1) {
        for (int j = kk >> 1; j > 0; j >>= 1) {
            for (int q = tid; q < SZ; q += kB) {
                const int p = q ^ j;
                if (p > q) {
                    const int x = key[q], y = key[p];
                    const bool up = (q & kk) == 0;
                    if ((x > y) == up) {
                        key[q] = y;
                        key[p] = x;
                    }
                }
            }
            __syncthreads();
        }
    }
    int cnt = 0;
    for (int q = tid * C; q < tid * C + C; ++q) {
        const int v = key[q];
        if (v != kEmpty && (q == 0 || key[q - 1] != v)) ++cnt;
    }
    // exclusive scan over the wave
    int x = cnt;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const int y = __shfl_up(x, d, 64);
        if (tid >= d) x += y;
    }
    const int total = __shfl(x, 63, 64);
    int pos = x - cnt;
    if (total > apply_detail::kWaveChunks) {
        if (tid == 0) wnu[t] = -1;
        return;  // whole wave
    }
    for (int q = tid * C; q < tid * C + C; ++q) {
        const int v = key[q];
        if (v != kEmpty && (q == 0 || key[q - 1] != v)) {
            uniq[pos] = v;
            wcol[t * apply_detail::kWaveChunks + pos] = v * CW;
            ++pos;
        }
    }
    __syncthreads();
    for (int e = 0; e < n; ++e) {
        const int col = col32[b + e];
        const int c = col / CW;
        int lo = 0, hi = total;
        while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            if (uniq[mid] < c)
                lo = mid + 1;
            else
                hi = mid;
        }
        lpos[b + e] = static_cast<std::uint16_t>(lo * CW + col % CW);
    }
    if (tid == 0) wnu[t] = total;
}

template <int E>
void launch_wave_build(int cw, unsigned g, const InterpolationMatrix &A, std::int32_t *wnu, std::int32_t *wcol, std::uint16_t *lpos, hipStream_t s) {
    constexpr int B = apply_detail::kWaveSeg;
    if (cw == 2)
        wave_tile_build_kernel<E, 2><<<g, B, 0, s>>>(A.seg_begin.get(), A.seg_len.get(), A.n_seg, A.col32.get(), wnu, wcol, lpos);
    else if (cw == 4)
        wave_tile_build_kernel<E, 4><<<g, B, 0, s>>>(A.seg_begin.get(), A.seg_len.get(), A.n_seg, A.col32.get(), wnu, wcol, lpos);
    else
        wave_tile_build_kernel<E, 8><<<g, B, 0, s>>>(A.seg_begin.get(), A.seg_len.get(), A.n_seg, A.col32.get(), wnu, wcol, lpos);
    AXIS_LAUNCH_CHECK();
}

}  // namespace

void InterpolationMatrix::prepare_wave_tiles(int cw, hipStream_t s) {
    const int ws = apply_detail::width_slot(cw);
    if (wave_ready[ws]) return;
    prepare_apply(s);
    PhaseTrace tr("prepare_wave_tiles", s);
    n_wtiles = (n_seg + apply_detail::kWaveSeg - 1) / apply_detail::kWaveSeg;
    const std::size_t nw = static_cast<std::size_t>(std::max<index_t>(n_wtiles, 1));
    wnu[ws].resize(nw);
    wcol[ws].resize(nw * apply_detail::kWaveChunks);
    wlpos[ws].resize(static_cast<std::size_t>(std::max<index_t>(nnz, 1)));
    if (n_wtiles > 0) {
        const unsigned g = static_cast<unsigned>(n_wtiles);
        switch (seg_cap) {
            case 4: launch_wave_build<4>(cw, g, *this, wnu[ws].get(), wcol[ws].get(), wlpos[ws].get(), s); break;
            case 8: launch_wave_build<8>(cw, g, *this, wnu[ws].get(), wcol[ws].get(), wlpos[ws].get(), s); break;
            case 12: launch_wave_build<12>(cw, g, *this, wnu[ws].get(), wcol[ws].get(), wlpos[ws].get(), s); break;
            default: launch_wave_build<16>(cw, g, *this, wnu[ws].get(), wcol[ws].get(), wlpos[ws].get(), s); break;
        }
    }
    AXIS_HIP_CHECK(hipStreamSynchronize(s));
    tr.mark("wave_tiles");
    wave_ready[ws] = true;
}

int InterpolationMatrix::row_length_hist(hipStream_t s) {
    constexpr int kH = apply_detail::kMaxE + 2;
    static_assert(kH == sizeof(len_hist) / sizeof(len_hist[0]), "histogram bins");
    if (!len_hist_ready) {
        DeviceBuffer<index_t> d_hist(kH);
        AXIS_HIP_CHECK(hipMemsetAsync(d_hist.get(), 0, sizeof(index_t) * kH, s));
        if (n_dst > 0) {
            len_hist_kernel<<<grid_for(n_dst, 256), 256, 0, s>>>(row_ptr.get(), n_dst, d_hist.get());
            AXIS_LAUNCH_CHECK();
        }
        d_hist.download(len_hist, kH, s);
        AXIS_HIP_CHECK(hipStreamSynchronize(s));
        len_hist_ready = true;
    }
    // E = the smallest register segment (4, 8, 12, 16) that leaves at most
    // 0.1% of the rows longer than one segment
    for (int cap : {4, 8, 12, 16}) {
        index_t above = 0;
        for (int l = cap + 1; l < kH; ++l) above += len_hist[l];
        if (above * 1000 <= n_dst) return cap;
    }
    return apply_detail::kMaxE;
}

void InterpolationMatrix::prepare_apply(hipStream_t s) {
    if (apply_ready) return;
    PhaseTrace tr("prepare_apply", s);
    if (n_src >= (index_t(1) << 31) - 1) throw InvalidArgument("apply: more than 2^31 - 2 source cells");
    if (n_dst >= (index_t(1) << 31)) throw InvalidArgument("apply: more than 2^31 destination cells");
    const index_t n1 = std::max<index_t>(nnz, 1);
    col32.resize(n1);
    val_hi.resize(n1);
    val_lo.resize(n1);
    if (nnz > 0) {
        split_kernel<<<grid_for(nnz, 256), 256, 0, s>>>(val.get(), col.get(), nnz, val_hi.get(), val_lo.get(), col32.get());
        AXIS_LAUNCH_CHECK();
    }
    seg_cap = row_length_hist(s);
    tr.mark("split_hist");
    // Processing order: rows grouped by a coarse spatial cell of their
    // centroid (about 256 destination cells per cell), index order inside a
    // cell.  A tile of 256 segments then covers a compact patch, whose source
    // columns are few and shared, and writes Y in runs of neighbouring indices.
    const index_t nd1 = std::max<index_t>(n_dst, 1);
    DeviceBuffer<std::uint64_t> key(nd1);
    int coarse_bits = 0;
    if (row_key.size() >= static_cast<std::size_t>(nd1) && n_dst > 0) {
        // cells of side 2.2 / 2^b on the [-1.1, 1.1]^3 Morton grid holding ~256 unit-sphere cells
        const double side = std::sqrt(4.0 * kPi * 256.0 / static_cast<double>(n_dst));
        coarse_bits = static_cast<int>(std::lround(std::log2(2.2 / side)));
        coarse_bits = std::clamp(coarse_bits, 1, 10);
    }
    if (n_dst > 0) {
        order_key_kernel<<<grid_for(n_dst, 256), 256, 0, s>>>(coarse_bits ? row_key.get() : nullptr, 3 * (10 - coarse_bits), n_dst, key.get());
        AXIS_LAUNCH_CHECK();
        // the low 32 bits (the row) are already ascending: sort the cell bits only
        if (coarse_bits) detail::radix_sort_keys_u64(key.get(), static_cast<std::size_t>(n_dst), s, 32 + 3 * coarse_bits, 32);
    }
    tr.mark("order_sort");
    DeviceBuffer<index_t> c_seg(nd1), c_slot(nd1), c_long(nd1), o_seg(nd1 + 1), o_slot(nd1 + 1), o_long(nd1 + 1);
    if (n_dst > 0) {
        seg_count_kernel<<<grid_for(n_dst, 256), 256, 0, s>>>(key.get(), row_ptr.get(), n_dst, seg_cap, c_seg.get(), c_slot.get(), c_long.get());
        AXIS_LAUNCH_CHECK();
    }
    detail::exclusive_scan_i64(c_seg.get(), o_seg.get(), n_dst, s);
    detail::exclusive_scan_i64(c_slot.get(), o_slot.get(), n_dst, s);
    detail::exclusive_scan_i64(c_long.get(), o_long.get(), n_dst, s);
    index_t tot[3] = {0, 0, 0};
    AXIS_HIP_CHECK(hipMemcpyAsync(&tot[0], o_seg.get() + n_dst, sizeof(index_t), hipMemcpyDeviceToHost, s));
    AXIS_HIP_CHECK(hipMemcpyAsync(&tot[1], o_slot.get() + n_dst, sizeof(index_t), hipMemcpyDeviceToHost, s));
    AXIS_HIP_CHECK(hipMemcpyAsync(&tot[2], o_long.get() + n_dst, sizeof(index_t), hipMemcpyDeviceToHost, s));
    AXIS_HIP_CHECK(hipStreamSynchronize(s));
    if (tot[1] >= (index_t(1) << 31)) throw InvalidArgument("apply: too many long-row segments");
    n_seg = tot[0];
    n_slots = static_cast<int>(tot[1]);
    n_long_rows = static_cast<int>(tot[2]);
    const std::size_t ns1 = static_cast<std::size_t>(std::max<index_t>(n_seg, 1));
    seg_row.resize(ns1);
    seg_begin.resize(ns1);
    seg_len.resize(ns1);
    seg_slot.resize(ns1);
    long_rows.resize(static_cast<std::size_t>(n_long_rows) + 1);
    long_slot.resize(static_cast<std::size_t>(n_long_rows) + 1);
    if (n_dst > 0) {
        seg_fill_kernel<<<grid_for(n_dst, 256), 256, 0, s>>>(key.get(), row_ptr.get(), n_dst, seg_cap, o_seg.get(), o_slot.get(), o_long.get(),
                                                            seg_row.get(), seg_begin.get(), seg_len.get(), seg_slot.get(), long_rows.get(),
                                                            long_slot.get());
        AXIS_LAUNCH_CHECK();
    }
    const std::int32_t end_slot = n_slots;
    AXIS_HIP_CHECK(hipMemcpyAsync(long_slot.get() + n_long_rows, &end_slot, sizeof(std::int32_t), hipMemcpyHostToDevice, s));
    tr.mark("segments");
    n_tiles = (n_seg + apply_detail::kTileSeg - 1) / apply_detail::kTileSeg;
    tiles_ready = false;
    for (bool &w : wave_ready) w = false;
    AXIS_HIP_CHECK(hipStreamSynchronize(s));
    apply_ready = true;
}

void InterpolationMatrix::prepare_tiles(hipStream_t s) {
    if (tiles_ready) return;
    prepare_apply(s);
    PhaseTrace tr("prepare_tiles", s);
    const index_t n1 = std::max<index_t>(nnz, 1);
    tile_nu.resize(static_cast<std::size_t>(std::max<index_t>(n_tiles, 1)));
    tile_col.resize(static_cast<std::size_t>(std::max<index_t>(n_tiles, 1)) * apply_detail::kTileChunks);
    t_lcol.resize(static_cast<std::size_t>(n1));
    if (n_tiles > 0) {
        const unsigned g = static_cast<unsigned>(n_tiles);
        const int B = apply_detail::kTileSeg;
        switch (seg_cap) {
            case 4: tile_build_kernel<4><<<g, B, 0, s>>>(seg_begin.get(), seg_len.get(), n_seg, col32.get(), tile_nu.get(), tile_col.get(), t_lcol.get()); break;
            case 8: tile_build_kernel<8><<<g, B, 0, s>>>(seg_begin.get(), seg_len.get(), n_seg, col32.get(), tile_nu.get(), tile_col.get(), t_lcol.get()); break;
            case 12: tile_build_kernel<12><<<g, B, 0, s>>>(seg_begin.get(), seg_len.get(), n_seg, col32.get(), tile_nu.get(), tile_col.get(), t_lcol.get()); break;
            default: tile_build_kernel<16><<<g, B, 0, s>>>(seg_begin.get(), seg_len.get(), n_seg, col32.get(), tile_nu.get(), tile_col.get(), t_lcol.get()); break;
        }
        AXIS_LAUNCH_CHECK();
    }
    AXIS_HIP_CHECK(hipStreamSynchronize(s));
    tr.mark("tiles");
    tiles_ready = true;
}

// plain applies of at most this many fields through a matrix without an
// apply layout take the direct kernels
constexpr index_t kDirectMaxFields = 2;

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
    if (mode != ApplyMode::Plain && (xt == FieldType::F64 || yt == FieldType::BF16))
        throw InvalidArgument("batch_apply: the mixed apply takes FP32/BF16 fields and writes FP32/FP64");
    if (k == 0 || n_dst == 0) return;
    std::lock_guard<std::mutex> lock(A.apply_mu);
    if (apply_detail::capturing(s)) {
        // recorded into a graph: nothing may be built, allocated or synchronised
        const apply_detail::Launch L = apply_detail::choose(A, k, xt, yt, mode);
        if (!apply_detail::layout_ready(A, L, xt, x))
            throw InvalidArgument("batch_apply: a stream under graph capture needs the matrix's apply layout from an earlier apply with these field types");
        apply_detail::launch(A, x, xt, y, yt, k, mode, L, s);
        return;
    }
    if (mode == ApplyMode::Plain && !A.apply_ready && k <= kDirectMaxFields && n_dst < (index_t(1) << 31)) {
        // a few fields through a matrix without an apply layout: the direct
        // kernels give the segmented apply's bits without building the layout
        const int E = A.row_length_hist(s);
        index_t n_long = 0;
        for (int l = E + 1; l < apply_detail::kMaxE + 2; ++l) n_long += A.len_hist[l];
        DeviceBuffer<std::int32_t> list(std::max<index_t>(n_long, 1));
        DeviceBuffer<int> cnt(1);
        AXIS_HIP_CHECK(hipMemsetAsync(cnt.get(), 0, sizeof(int), s));
        auto rows = [&](auto kern) {
            kern<<<grid_for(n_dst, 256), 256, 0, s>>>(A.row_ptr.get(), A.col.get(), A.val.get(), n_dst, n_src, static_cast<const double *>(x),
                                                      static_cast<double *>(y), k, list.get(), cnt.get());
            AXIS_LAUNCH_CHECK();
        };
        switch (E) {
            case 4: rows(direct_rows_kernel<4>); break;
            case 8: rows(direct_rows_kernel<8>); break;
            case 12: rows(direct_rows_kernel<12>); break;
            case 16: rows(direct_rows_kernel<16>); break;
            default: throw RuntimeError("batch_apply: unexpected segment length " + std::to_string(E));
        }
        if (n_long > 0) {
            direct_long_kernel<<<static_cast<unsigned>((n_long + 3) / 4), 256, 0, s>>>(A.row_ptr.get(), A.col.get(), A.val.get(), n_dst, n_src,
                                                                                        static_cast<const double *>(x), static_cast<double *>(y), k, E,
                                                                                        list.get(), static_cast<int>(n_long));
            AXIS_LAUNCH_CHECK();
        }
        // (list and cnt go back to the block cache, which hands a block out
        // again only after a device synchronisation: no wait here, so the
        // async entry point stays asynchronous)
        return;
    }
    A.prepare_apply(s);
    apply_detail::Launch L = apply_detail::choose(A, k, xt, yt, mode);
    apply_detail::launch(A, x, xt, y, yt, k, mode, L, s);
}

// ------------------------------------------------------ field_view entry points
namespace {
[[noreturn]] void extent_error(const char *fn, const char *what, std::size_t got, const char *whose, index_t want) {
    throw InvalidArgument(std::string("axis::solver::") + fn + ": " + what + " (" + std::to_string(got) + ") != " + whose + " (" +
                          std::to_string(want) + ")");
}
}  // namespace

void apply(InterpolationMatrix &A, field_view<const double, 1> src, field_view<double, 1> dst, hipStream_t s) {
    // apply.hpp:63-71: validation before dst is touched
    if (src.extent(0) != static_cast<std::size_t>(A.n_src)) extent_error("apply", "src.extent(0)", src.extent(0), "matrix.n_src()", A.n_src);
    if (dst.extent(0) != static_cast<std::size_t>(A.n_dst)) extent_error("apply", "dst.extent(0)", dst.extent(0), "matrix.n_dst()", A.n_dst);
    batch_apply(A, src.data_handle(), FieldType::F64, dst.data_handle(), FieldType::F64, A.n_src, A.n_dst, 1, ApplyMode::Plain, s);
    AXIS_HIP_CHECK(hipStreamSynchronize(s));
}

void apply(InterpolationMatrix &A, field_view<const double, 1> src, field_view<double, 1> dst, field_view<const std::int32_t, 1> dst_mask,
           hipStream_t s) {
    if (src.extent(0) != static_cast<std::size_t>(A.n_src))
        extent_error("apply (masked)", "src.extent(0)", src.extent(0), "matrix.n_src()", A.n_src);
    if (dst.extent(0) != static_cast<std::size_t>(A.n_dst))
        extent_error("apply (masked)", "dst.extent(0)", dst.extent(0), "matrix.n_dst()", A.n_dst);
    masked_apply(A, src.data_handle(), dst.data_handle(), dst_mask.data_handle(), static_cast<index_t>(dst_mask.extent(0)), A.n_src, A.n_dst, 1, s);
    AXIS_HIP_CHECK(hipStreamSynchronize(s));
}

void batch_apply(InterpolationMatrix &A, field_view<const double, 2> src, field_view<double, 2> dst, hipStream_t s) {
    // apply.hpp:280-292
    if (src.extent(0) != static_cast<std::size_t>(A.n_src))
        extent_error("batch_apply", "src.extent(0)", src.extent(0), "matrix.n_src()", A.n_src);
    if (dst.extent(0) != static_cast<std::size_t>(A.n_dst))
        extent_error("batch_apply", "dst.extent(0)", dst.extent(0), "matrix.n_dst()", A.n_dst);
    if (src.extent(1) != dst.extent(1))
        extent_error("batch_apply", "src.extent(1)", src.extent(1), "dst.extent(1)", static_cast<index_t>(dst.extent(1)));
    batch_apply(A, src.data_handle(), FieldType::F64, dst.data_handle(), FieldType::F64, A.n_src, A.n_dst, static_cast<index_t>(src.extent(1)),
                ApplyMode::Plain, s);
    AXIS_HIP_CHECK(hipStreamSynchronize(s));
}

void nan_batch_apply(InterpolationMatrix &A, field_view<const double, 2> src, field_view<double, 2> dst, double na_thres, hipStream_t s) {
    // apply.hpp:363-371
    if (src.extent(0) != static_cast<std::size_t>(A.n_src)) throw InvalidArgument("axis::solver::nan_batch_apply: src.extent(0) != matrix.n_src()");
    if (dst.extent(0) != static_cast<std::size_t>(A.n_dst)) throw InvalidArgument("axis::solver::nan_batch_apply: dst.extent(0) != matrix.n_dst()");
    if (src.extent(1) != dst.extent(1)) throw InvalidArgument("axis::solver::nan_batch_apply: src.extent(1) != dst.extent(1)");
    nan_batch_apply(A, src.data_handle(), dst.data_handle(), A.n_src, A.n_dst, static_cast<index_t>(src.extent(1)), na_thres, s);
    AXIS_HIP_CHECK(hipStreamSynchronize(s));
}

// ------------------------------------------------ apply without allocating
namespace {

// Long-row partial sums of batch_apply_noalloc: one array per GPU the library
// runs on, held until the process ends.  Host code must use the per-device
// address from hipGetSymbolAddress, never the array's name.
constexpr std::size_t kNoAllocScratchDoubles = std::size_t(1) << 22;  // 32 MiB
static __device__ __align__(256) double scratch[kNoAllocScratchDoubles];

double *scratch_on_current_device() {
    static std::mutex mu;
    static std::map<int, double *> ptrs;
    int dev = 0;
    AXIS_HIP_CHECK(hipGetDevice(&dev));
    std::lock_guard<std::mutex> g(mu);
    auto it = ptrs.find(dev);
    if (it != ptrs.end()) return it->second;
    double *p = nullptr;
    AXIS_HIP_CHECK(hipGetSymbolAddress(reinterpret_cast<void **>(&p), HIP_SYMBOL(scratch)));
    ptrs[dev] = p;
    return p;
}

}  // namespace

std::size_t noalloc_scratch_doubles() { return kNoAllocScratchDoubles; }

void batch_apply_noalloc(InterpolationMatrix &A, const void *x, FieldType xt, void *y, FieldType yt, index_t n_src, index_t n_dst, index_t k,
                         ApplyMode mode, hipStream_t s) {
    if (n_src != A.n_src) throw InvalidArgument("axis::solver::batch_apply: src.extent(0) (" + std::to_string(n_src) + ") != matrix.n_src() (" +
                                                std::to_string(A.n_src) + ")");
    if (n_dst != A.n_dst) throw InvalidArgument("axis::solver::batch_apply: dst.extent(0) (" + std::to_string(n_dst) + ") != matrix.n_dst() (" +
                                                std::to_string(A.n_dst) + ")");
    if (k < 0) throw InvalidArgument("batch_apply: negative field count");
    if (mode == ApplyMode::Plain && (xt != FieldType::F64 || yt != FieldType::F64))
        throw InvalidArgument("batch_apply: the plain apply takes FP64 fields");
    if (mode != ApplyMode::Plain && (xt == FieldType::F64 || yt == FieldType::BF16))
        throw InvalidArgument("batch_apply: the mixed apply takes FP32/BF16 fields and writes FP32/FP64");
    if (k == 0 || n_dst == 0) return;
    std::lock_guard<std::mutex> lock(A.apply_mu);
    if (!A.apply_ready)
        throw InvalidArgument("batch_apply_noalloc: the matrix has no apply layout yet (one axis_batch_apply_c with these field types builds it)");
    const apply_detail::Launch L = apply_detail::choose(A, k, xt, yt, mode);
    if (!apply_detail::layout_ready(A, L, xt, x))
        throw InvalidArgument("batch_apply_noalloc: the matrix has no staged tiles yet (one axis_batch_apply_c with these field types builds them)");
    const std::size_t need = static_cast<std::size_t>(A.n_slots) * static_cast<std::size_t>(k);
    if (need > kNoAllocScratchDoubles)
        throw InvalidArgument("batch_apply_noalloc: " + std::to_string(A.n_slots) + " long-row partial sums x " + std::to_string(k) +
                              " fields exceed the scratch maximum of " + std::to_string(kNoAllocScratchDoubles) + " doubles");
    apply_detail::SpecialArgs sa;
    if (A.n_slots > 0) sa.partials = scratch_on_current_device();
    apply_detail::launch(A, x, xt, y, yt, k, mode, L, s, sa);
}

void set_apply_shape(InterpolationMatrix &A, index_t k, FieldType xt, FieldType yt, ApplyMode mode, const ApplyShape &shape) {
    if (k <= 0) throw InvalidArgument("set_apply_shape: no fields");
    if (shape.fields_per_step != 1 && shape.fields_per_step != 2 && shape.fields_per_step != 4)
        throw InvalidArgument("set_apply_shape: fields per step must be 1, 2 or 4");
    if (shape.field_chunks < 1 || shape.field_chunks > 1024) throw InvalidArgument("set_apply_shape: field chunks must be in [1, 1024]");
    apply_detail::Launch L;
    L.fu = shape.fields_per_step;
    L.chunks = shape.field_chunks;
    L.tiled = shape.staged && !shape.wave;
    L.wave = shape.wave;
    std::lock_guard<std::mutex> g(apply_detail::tune_mutex());
    apply_detail::tune_cache()[apply_detail::TuneKey{A.n_dst, A.nnz, k, static_cast<int>(xt), static_cast<int>(yt), static_cast<int>(mode)}] = L;
}

ApplyShape autotune_apply(InterpolationMatrix &A, const void *x, FieldType xt, void *y, FieldType yt, index_t k, ApplyMode mode, hipStream_t s) {
    if (k <= 0) throw InvalidArgument("autotune_apply: no fields");
    std::lock_guard<std::mutex> lock(A.apply_mu);
    A.prepare_apply(s);
    hipEvent_t e0, e1;
    AXIS_HIP_CHECK(hipEventCreate(&e0));
    AXIS_HIP_CHECK(hipEventCreate(&e1));
    apply_detail::Launch best = apply_detail::heuristic(A, k, mode);
    float best_ms = 1e30f;
    for (int kind : {0, 1, 2})  // direct gathers, staged tiles, wave tiles
    for (int fu : {1, 2, 4}) {
        if (fu > 1 && fu / 2 >= k) continue;
        for (int chunks : {1, 2, 4, 8, 16, 32, 64}) {
            if (chunks > 1 && (chunks / 2) * fu >= k) continue;
            apply_detail::Launch L;
            L.fu = fu;
            L.chunks = chunks;
            L.tiled = kind == 1;
            L.wave = kind == 2;
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
    return ApplyShape{best.fu, best.chunks, best.tiled || best.wave, best.wave};
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

// ------------------------------------------- NaN-aware and masked applies
// Both run through the segment/tile machinery of the apply (spatial order,
// LDS-staged source columns) with their own row accumulators; long rows are
// finished whole by a per-(row, field) kernel, so every row is summed in CSR
// order.
namespace {
apply_detail::Launch special_launch(InterpolationMatrix &A, index_t k, hipStream_t s) {
    A.prepare_apply(s);
    apply_detail::Launch L = apply_detail::heuristic(A, k, ApplyMode::Plain);
    L.tiled = true;
    A.prepare_tiles(s);
    return L;
}
}  // namespace

void nan_batch_apply(InterpolationMatrix &A, const double *x, double *y, index_t n_src, index_t n_dst, index_t k, double na_thres,
                     hipStream_t s) {
    // apply.hpp:362-370: validation before dst is touched
    if (n_src != A.n_src) throw InvalidArgument("axis::solver::nan_batch_apply: src.extent(0) != matrix.n_src()");
    if (n_dst != A.n_dst) throw InvalidArgument("axis::solver::nan_batch_apply: dst.extent(0) != matrix.n_dst()");
    if (k < 0) throw InvalidArgument("axis::solver::nan_batch_apply: src.extent(1) != dst.extent(1)");
    if (k == 0 || n_dst == 0) return;
    std::lock_guard<std::mutex> lock(A.apply_mu);
    const apply_detail::Launch L = special_launch(A, k, s);
    apply_detail::SpecialArgs sa;
    sa.thres = na_thres;
    apply_detail::launch_xy<apply_detail::kNanAware, double, double>(A, x, y, k, L, s, sa);
}

void masked_apply(InterpolationMatrix &A, const double *x, double *y, const std::int32_t *dst_mask, index_t n_mask, index_t n_src,
                  index_t n_dst, index_t k, hipStream_t s) {
    // apply.hpp:453-461
    if (n_src != A.n_src)
        throw InvalidArgument("axis::solver::apply (masked): src.extent(0) (" + std::to_string(n_src) + ") != matrix.n_src() (" +
                              std::to_string(A.n_src) + ")");
    if (n_dst != A.n_dst)
        throw InvalidArgument("axis::solver::apply (masked): dst.extent(0) (" + std::to_string(n_dst) + ") != matrix.n_dst() (" +
                              std::to_string(A.n_dst) + ")");
    if (k < 0) throw InvalidArgument("apply (masked): negative field count");
    if (k == 0 || n_dst == 0) return;
    std::lock_guard<std::mutex> lock(A.apply_mu);
    const apply_detail::Launch L = special_launch(A, k, s);
    apply_detail::SpecialArgs sa;
    sa.mask = (dst_mask && n_mask == n_dst) ? dst_mask : nullptr;  // apply.hpp:464
    apply_detail::launch_xy<apply_detail::kMasked, double, double>(A, x, y, k, L, s, sa);
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
