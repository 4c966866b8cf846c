// apply_kernels.hpp -- device code of the sparse apply Y = A X on layout_left
// fields, shared by the single-matrix launch (apply.hip) and the persistent
// multi-pair batch (batch.hip).
//
// Design (HBM-bound: touch every byte of A, X and Y about once):
//   * the CSR rows are cut into "segments" of at most E nonzeros (E = 4, 8,
//     12 or 16, chosen per matrix so that almost every row is one segment)
//     and ordered spatially (destination cells grouped by a coarse Morton
//     cell of their centroid);
//   * one lane per segment: its (col, w) entries are loaded once into
//     registers and reused while the lane streams through a chunk of fields,
//     FU fields per step.  A lane whose segment is a whole row writes Y
//     directly;
//   * segments of a long row (e.g. a polar cell with ~1600 nonzeros) write
//     partial sums to a workspace; a second kernel adds each long row's
//     partials in segment order (deterministic) and writes Y;
//   * staged tiles (spmm_tile_kernel): a workgroup's 256 segments form a
//     tile whose distinct source columns (a compact patch, ~0.3 per nonzero)
//     are read once per field, in sorted order, into LDS; the lanes then
//     gather from LDS.  Inside a tile the segments are sorted by length, so
//     the lanes of a wave run the same number of terms;
//   * work items are (tile, field chunk); the blockIdx -> item map gives each
//     XCD a contiguous run of tiles per chunk, so neighbouring tiles, which
//     share source columns, share that XCD's L2.
//
// Numerics (template parameter NUM):
//   kPlain   FP64 fields, FP64 weights, FP64 FMA chain (the plain method);
//   kComp32  FP32/BF16 fields, weights split into FP32 hi + lo, FP32
//            accumulation with TwoProduct/TwoSum compensation;
//   kAcc64   FP32/BF16 fields, FP64 weights, FP64 FMA accumulation (on this
//            GPU a FP64 FMA issues at the rate of an unpacked FP32 FMA, so
//            this is both the cheapest and the most accurate reduction of
//            narrow fields).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <map>
#include <mutex>
#include <tuple>
#include <type_traits>

#include "axis_hip.hpp"

namespace axis_hip::apply_detail {

constexpr int kBlock = 256;        // segments per workgroup of the direct-gather kernel
constexpr int kTileSeg = 256;      // segments per staged tile (= threads of its workgroup; 512 measured slower)
constexpr int kMaxE = 16;          // largest register segment
constexpr int kChunk = 4;          // source columns staged per aligned chunk (one 16-byte FP32 load)
constexpr int kTileChunks = kTileSeg;             // chunks a staged tile holds: one per lane
constexpr int kTileCols = kTileChunks * kChunk;   // staged values per field

enum Num : int { kPlain = 0, kComp32 = 1, kAcc64 = 2, kNanAware = 3, kMasked = 4 };

/// kNanAware / kMasked: FP64 fields and weights, each whole row summed in CSR
/// order with separate multiply and add (the reference's host arithmetic of
/// nan_batch_apply and the masked apply, apply.hpp:359-484); long rows are
/// finished by long_rows_special_kernel instead of partial sums.
constexpr bool is_special(int num) { return num == kNanAware || num == kMasked; }

/// Launch shape chosen by the autotuner: FU fields per step and the number of
/// field chunks the k fields are split into (more chunks = more parallelism,
/// more re-reads of A).
struct Launch {
    int fu = 4;
    int chunks = 1;
    bool tiled = false;  // stage each tile's source values in LDS
    bool wave = false;   // wave tiles: each wave stages its own 64 segments' source chunks (no workgroup barriers)
};

// Wave tiles: the segments [64 w, 64 w + 64) of the processing order form wave
// tile w.  Its distinct aligned 16-byte chunks of a field (kWaveNQ per lane at
// most) are listed by first column in wcol[w * kWaveChunks ...]; per nonzero,
// wlpos holds its column's element position in the staged chunks.  The chunk
// width depends on the field type (16 bytes: 2 FP64, 4 FP32, 8 BF16 columns),
// so the layout is built per width.  wnu[w] = -1: the wave gathers directly.
constexpr int kWaveSeg = 64;
constexpr int kWaveNQ = 2;
constexpr int kWaveChunks = kWaveSeg * kWaveNQ;
constexpr int kWaveStage = kWaveChunks * 16;  // bytes of one field's staged chunks
template <class XT>
constexpr int chunk_cols() {
    return 16 / static_cast<int>(sizeof(XT));
}
/// index of a chunk width (2, 4, 8 columns) in the per-width layouts
constexpr int width_slot(int cw) { return cw == 2 ? 0 : (cw == 4 ? 1 : 2); }

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
    __device__ inline double value() const { return static_cast<double>(s) + static_cast<double>(c); }
};

/// Operands of one apply.
struct Mat {
    const std::int32_t *col;
    const double *val;
    const float *hi, *lo;
    index_t n_src, n_dst;
    // segments, in processing order
    const std::int32_t *seg_row;   // destination row
    const index_t *seg_begin;      // first nonzero
    const std::int32_t *seg_len;   // <= E
    const std::int32_t *seg_slot;  // -1: whole row; else partial-sum slot
    index_t n_seg;
    // long rows: partial slots [long_slot[j], long_slot[j + 1]) of row long_row[j]
    const std::int32_t *long_row, *long_slot;
    int n_long, n_slots;
    // tiles (segments [256 t, 256 t + 256)): number of distinct aligned
    // 4-column chunks their nonzeros touch (-1: too many, gather directly),
    // the chunks' first columns, and per nonzero the position of its column
    // in the tile's staged values (4 * chunk + column % 4)
    const std::int32_t *tile_nu, *tile_col;
    const std::uint16_t *lcol;
    index_t n_tiles;
    // nan-aware / masked applies: the valid-fraction threshold, the
    // destination mask (nullptr: none)
    double thres;
    const std::int32_t *mask;
    // wave tiles of the launch's field type (see kWaveSeg)
    const std::int32_t *wnu, *wcol;
    const std::uint16_t *wlpos;
    index_t n_wt;
};

/// Hardware places workgroup b on XCD b % 8; give XCD x a contiguous range of
/// work items.
__device__ inline unsigned xcd_remap(unsigned bid, unsigned n) {
    const unsigned xcd = bid & 7u, idx = bid >> 3;
    const unsigned q = n >> 3, rem = n & 7u;
    return xcd * q + (xcd < rem ? xcd : rem) + idx;
}

template <class XT>
__device__ inline float ldv(const XT *p) {
    return ld_f32(p);
}
__device__ inline double ldv(const double *p) { return *p; }

template <int NUM>
struct NumTraits {
    using V = float;   // field value held in registers / LDS
    using W = double;  // weight held in registers
};
template <>
struct NumTraits<kPlain> {
    using V = double;
    using W = double;
};
template <>
struct NumTraits<kComp32> {
    using V = float;
    using W = float;
};
template <>
struct NumTraits<kNanAware> {
    using V = double;
    using W = double;
};
template <>
struct NumTraits<kMasked> {
    using V = double;
    using W = double;
};

/// Row-sum accumulator of one numerics mode.
template <int NUM>
struct RowAcc;
template <>
struct RowAcc<kPlain> {
    double s = 0.0;
    __device__ inline void add(double h, double, double x) { s = fma(h, x, s); }
    __device__ inline double value() const { return s; }
};
template <>
struct RowAcc<kAcc64> {
    double s = 0.0;
    __device__ inline void add(double h, double, float x) { s = fma(h, static_cast<double>(x), s); }
    __device__ inline double value() const { return s; }
};
template <>
struct RowAcc<kComp32> : Acc32 {};
template <>
struct RowAcc<kNanAware> {
    double ws = 0.0, sv = 0.0, total = 0.0, thres = 0.0;
    __device__ inline void add(double h, double, double x) {
        if (!isnan(x)) {
            ws += h * x;
            sv += h;
        }
    }
    __device__ inline double value() const {
        if (sv == 0.0 || total == 0.0) return __builtin_nan("");
        const double frac = sv / total;
        return frac < (1.0 - thres - 1e-6) ? __builtin_nan("") : ws / sv;
    }
};
template <>
struct RowAcc<kMasked> {
    double s = 0.0;
    __device__ inline void add(double h, double, double x) { s += h * x; }
    __device__ inline double value() const { return s; }
};

/// Start of a row's sums: the NaN-aware accumulator needs the row's total
/// weight (CSR order) and the threshold.
template <int NUM, int E, class W>
__device__ inline void begin_row(RowAcc<NUM> &acc, const Mat &A, const W (&h)[E], int n) {
    if constexpr (NUM == kNanAware) {
        double t = 0.0;
#pragma unroll
        for (int e = 0; e < E; ++e)
            if (e < n) t += h[e];
        acc.total = t;
        acc.thres = A.thres;
    }
}

/// Whether the special modes write row r (whole rows only; masked rows stay).
__device__ inline bool special_writes(const Mat &A, index_t r, int slot) { return slot < 0 && (A.mask == nullptr || A.mask[r] != 0); }

template <int NUM>
__device__ inline void load_w(const Mat &A, index_t q, typename NumTraits<NUM>::W &h, typename NumTraits<NUM>::W &l) {
    if constexpr (NUM == kComp32) {
        h = A.hi[q];
        l = A.lo[q];
    } else {
        h = A.val[q];
        l = 0.0;
    }
}

/// Segment i of the processing order, fields [f0, f1) in steps of FU, source
/// values gathered straight from X.  Partial sums of long-row segments go to
/// P[f * n_slots + slot] (FP64).
template <int E, int FU, int NUM, class XT, class YT>
__device__ inline void segment_apply(const Mat &A, const XT *X, YT *Y, double *P, index_t i, index_t f0, index_t f1) {
    using V = typename NumTraits<NUM>::V;
    using W = typename NumTraits<NUM>::W;
    const index_t r = A.seg_row[i];
    const index_t b = A.seg_begin[i];
    const int n = A.seg_len[i];
    const int slot = A.seg_slot[i];
    std::int32_t c[E];
    W h[E], l[E];
#pragma unroll
    for (int e = 0; e < E; ++e) {
        c[e] = 0;
        h[e] = W(0);
        l[e] = W(0);
        if (e < n) {
            c[e] = A.col[b + e];
            load_w<NUM>(A, b + e, h[e], l[e]);
        }
    }
    for (index_t f = f0; f < f1; f += FU) {
        V x[FU][E];
#pragma unroll
        for (int u = 0; u < FU; ++u) {
            const index_t fu = (f + u < f1) ? f + u : f1 - 1;
            const XT *xf = X + fu * A.n_src;
#pragma unroll
            for (int e = 0; e < E; ++e) x[u][e] = (e < n) ? ldv(xf + c[e]) : V(0);
        }
#pragma unroll
        for (int u = 0; u < FU; ++u) {
            RowAcc<NUM> acc;
            begin_row(acc, A, h, n);
#pragma unroll
            for (int e = 0; e < E; ++e)
                if (e < n) acc.add(h[e], l[e], x[u][e]);
            const double v = acc.value();
            if (f + u < f1) {
                if constexpr (is_special(NUM)) {
                    if (special_writes(A, r, slot)) st(Y + (f + u) * A.n_dst + r, v);
                } else if (slot < 0) {
                    st(Y + (f + u) * A.n_dst + r, v);
                } else {
                    P[(f + u) * A.n_slots + slot] = v;
                }
            }
        }
    }
}

template <int E, int FU, int NUM, class XT, class YT>
__global__ void __launch_bounds__(kBlock) spmm_seg_kernel(Mat A, const XT *X, YT *Y, double *P, index_t k, unsigned n_blocks, int chunks) {
    const unsigned w = xcd_remap(blockIdx.x, gridDim.x);
    const unsigned sb = w % n_blocks, ch = w / n_blocks;
    const index_t i = static_cast<index_t>(sb) * kBlock + threadIdx.x;
    if (i >= A.n_seg) return;
    const index_t per = (k + chunks - 1) / chunks;
    const index_t f0 = static_cast<index_t>(ch) * per;
    const index_t f1 = f0 + per < k ? f0 + per : k;
    if (f0 >= f1) return;
    segment_apply<E, FU, NUM>(A, X, Y, P, i, f0, f1);
}

/// The 4 source values of an aligned chunk starting at column c0 of one
/// field: one 16-byte (FP32), 8-byte (BF16) or two 16-byte (FP64) loads when
/// the field is suitably aligned and the chunk lies inside it, else one load
/// per column (columns past n_src read as 0).
template <class V>
__device__ inline void load_chunk(const float *xf, index_t c0, index_t n_src, bool vec, V out[kChunk]) {
    if (vec && c0 + kChunk <= n_src) {
        const float4 v = *reinterpret_cast<const float4 *>(xf + c0);
        out[0] = v.x, out[1] = v.y, out[2] = v.z, out[3] = v.w;
    } else {
#pragma unroll
        for (int q = 0; q < kChunk; ++q) out[q] = c0 + q < n_src ? xf[c0 + q] : 0.0f;
    }
}
template <class V>
__device__ inline void load_chunk(const std::uint16_t *xf, index_t c0, index_t n_src, bool vec, V out[kChunk]) {
    if (vec && c0 + kChunk <= n_src) {
        const uint2 v = *reinterpret_cast<const uint2 *>(xf + c0);
        out[0] = __uint_as_float(v.x << 16), out[1] = __uint_as_float(v.x & 0xFFFF0000u);
        out[2] = __uint_as_float(v.y << 16), out[3] = __uint_as_float(v.y & 0xFFFF0000u);
    } else {
#pragma unroll
        for (int q = 0; q < kChunk; ++q) out[q] = c0 + q < n_src ? ld_f32(xf + c0 + q) : 0.0f;
    }
}
template <class V>
__device__ inline void load_chunk(const double *xf, index_t c0, index_t n_src, bool vec, V out[kChunk]) {
    if (vec && c0 + kChunk <= n_src) {
        const double2 a = *reinterpret_cast<const double2 *>(xf + c0), b = *reinterpret_cast<const double2 *>(xf + c0 + 2);
        out[0] = a.x, out[1] = a.y, out[2] = b.x, out[3] = b.y;
    } else {
#pragma unroll
        for (int q = 0; q < kChunk; ++q) out[q] = c0 + q < n_src ? xf[c0 + q] : 0.0;
    }
}

/// Tile t of segments, fields [f0, f1) (f0 < f1): the tile's distinct source
/// values of FU fields are loaded once into LDS (xs) by the whole workgroup
/// (each source value read from HBM once per tile, in sorted column order,
/// instead of once per nonzero), then every lane forms its segment's sums
/// from LDS.  Tiles whose columns do not fit fall back to direct gathers.
/// Called by every thread of the workgroup; returns after a barrier, so xs
/// may be reused at once.
/// LDS staging buffer of one tile workgroup: FU fields of the tile's chunks.
/// (A second buffer, one barrier per step, measured slower: the extra LDS
/// lowers occupancy more than the barrier costs.)
/// Slot kTileCols of each field holds 0: the padding terms of a lane whose
/// segment is shorter than its wave's longest read it (with zero weights), so
/// every lane runs the same terms without per-lane predication.
template <int FU, int NUM>
struct Stage {
    using V = typename NumTraits<NUM>::V;
    V xs[FU][kTileCols + kChunk];
};

/// The 4 values of an aligned chunk, unconditionally vector-loaded (fields
/// 16-byte aligned, n_src a multiple of 4, so the chunk lies inside the field).
template <class V>
__device__ inline void load_chunk_vec(const float *xf, V out[kChunk]) {
    const float4 v = *reinterpret_cast<const float4 *>(xf);
    out[0] = v.x, out[1] = v.y, out[2] = v.z, out[3] = v.w;
}
template <class V>
__device__ inline void load_chunk_vec(const std::uint16_t *xf, V out[kChunk]) {
    const uint2 v = *reinterpret_cast<const uint2 *>(xf);
    out[0] = __uint_as_float(v.x << 16), out[1] = __uint_as_float(v.x & 0xFFFF0000u);
    out[2] = __uint_as_float(v.y << 16), out[3] = __uint_as_float(v.y & 0xFFFF0000u);
}
template <class V>
__device__ inline void load_chunk_vec(const double *xf, V out[kChunk]) {
    const double2 a = *reinterpret_cast<const double2 *>(xf), b = *reinterpret_cast<const double2 *>(xf + 2);
    out[0] = a.x, out[1] = a.y, out[2] = b.x, out[3] = b.y;
}

template <int E, int FU, int NUM, class XT, class YT>
__device__ inline void tile_apply(const Mat &A, const XT *X, YT *Y, double *P, index_t t, index_t f0, index_t f1, Stage<FU, NUM> &S) {
    using V = typename NumTraits<NUM>::V;
    using W = typename NumTraits<NUM>::W;
    const index_t i = t * kTileSeg + threadIdx.x;
    const bool valid = i < A.n_seg;
    const int nu = A.tile_nu[t];
    if (nu < 0) {  // whole workgroup; one field per step keeps this rare path's registers low
        if (valid) segment_apply<E, 1, NUM>(A, X, Y, P, i, f0, f1);
        return;
    }
    index_t r = 0, b = 0;
    int n = 0, slot = -1;
    if (valid) {
        r = A.seg_row[i];
        b = A.seg_begin[i];
        n = A.seg_len[i];
        slot = A.seg_slot[i];
    }
    // terms run by this wave: its longest segment
    int n_max = n;
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) n_max = max(n_max, __shfl_xor(n_max, d, 64));
    const int n_wave = __builtin_amdgcn_readfirstlane(n_max);
    unsigned lc[E];  // byte offsets into one field's staged values (padding terms: the zero slot)
    W h[E], l[E];
#pragma unroll
    for (int e = 0; e < E; ++e) {
        lc[e] = static_cast<unsigned>(kTileCols) * static_cast<unsigned>(sizeof(V));
        h[e] = W(0);
        l[e] = W(0);
        if (e < n) {
            lc[e] = static_cast<unsigned>(A.lcol[b + e]) * static_cast<unsigned>(sizeof(V));
            load_w<NUM>(A, b + e, h[e], l[e]);
        }
    }
    if (threadIdx.x < FU) S.xs[threadIdx.x][kTileCols] = V(0);
    // this lane's chunk: columns c0 .. c0 + 3 of every field
    const int lane_chunk = static_cast<int>(threadIdx.x);
    const bool has = lane_chunk < nu;
    const index_t c0 = has ? static_cast<index_t>(A.tile_col[t * kTileChunks + lane_chunk]) : 0;
    // unconditional vector loads need 16-byte aligned fields holding whole chunks
    const bool vec = (A.n_src % kChunk) == 0 && (reinterpret_cast<std::uintptr_t>(X) % 16) == 0;
    auto sums = [&](int u, index_t f) {
        RowAcc<NUM> acc;
        begin_row(acc, A, h, n);
        const char *xb = reinterpret_cast<const char *>(&S.xs[u][0]);
        // terms in groups of 4 (E is a multiple of 4): the group's LDS reads
        // are issued together; the bound is wave-uniform
#pragma unroll
        for (int g = 0; g < E; g += 4) {
            if (g < n_wave) {
                V xv[4];
#pragma unroll
                for (int q = 0; q < 4; ++q) xv[q] = *reinterpret_cast<const V *>(xb + lc[g + q]);
#pragma unroll
                for (int q = 0; q < 4; ++q) acc.add(h[g + q], l[g + q], xv[q]);
            }
        }
        const double val = acc.value();
        if (valid && f < f1) {
            if constexpr (is_special(NUM)) {
                if (special_writes(A, r, slot)) st(Y + f * A.n_dst + r, val);
            } else if (slot < 0) {
                st(Y + f * A.n_dst + r, val);
            } else {
                P[f * A.n_slots + slot] = val;
            }
        }
    };
    if (!vec) {
        // per-column loads, one step of prefetch
        V v[FU][kChunk];
        auto fetch = [&](index_t f) {
#pragma unroll
            for (int u = 0; u < FU; ++u) {
                const index_t fu = (f + u < f1) ? f + u : f1 - 1;
                if (has) load_chunk<V>(X + fu * A.n_src, c0, A.n_src, false, v[u]);
            }
        };
        fetch(f0);
        for (index_t f = f0; f < f1; f += FU) {
            if (has) {
#pragma unroll
                for (int u = 0; u < FU; ++u)
#pragma unroll
                    for (int q = 0; q < kChunk; ++q) S.xs[u][lane_chunk * kChunk + q] = v[u][q];
            }
            __syncthreads();
            if (f + FU < f1) fetch(f + FU);
#pragma unroll
            for (int u = 0; u < FU; ++u) sums(u, f + u);
            __syncthreads();
        }
        return;
    }
    // software pipeline: the chunks of the next kDepth steps (FU fields each)
    // are in flight while a step's sums are formed.  Every lane loads (lanes
    // without a chunk load columns 0..3 and store them to slots no nonzero
    // reads), so no load sits under a per-lane branch.
    constexpr int kDepth = FU >= 4 ? 1 : 4 / FU;
    constexpr index_t kAhead = static_cast<index_t>(kDepth) * FU;
    V v[kDepth][FU][kChunk];
    auto fetch = [&](index_t f, V (&dst)[FU][kChunk]) {
#pragma unroll
        for (int u = 0; u < FU; ++u) {
            const index_t fu = (f + u < f1) ? f + u : f1 - 1;
            load_chunk_vec<V>(X + fu * A.n_src + c0, dst[u]);
        }
    };
#pragma unroll
    for (int d = 0; d < kDepth; ++d)
        if (f0 + d * FU < f1) fetch(f0 + d * FU, v[d]);
    for (index_t fb = f0; fb < f1; fb += kAhead) {
#pragma unroll
        for (int d = 0; d < kDepth; ++d) {
            const index_t f = fb + d * FU;
            if (f < f1) {  // uniform over the workgroup
#pragma unroll
                for (int u = 0; u < FU; ++u)
#pragma unroll
                    for (int q = 0; q < kChunk; ++q) S.xs[u][lane_chunk * kChunk + q] = v[d][u][q];
                __syncthreads();
                if (f + kAhead < f1) fetch(f + kAhead, v[d]);
#pragma unroll
                for (int u = 0; u < FU; ++u) sums(u, f + u);
                __syncthreads();
            }
        }
    }
}

template <int E, int FU, int NUM, class XT, class YT>
__global__ void __launch_bounds__(kTileSeg) spmm_tile_kernel(Mat A, const XT *X, YT *Y, double *P, index_t k, unsigned n_tiles, int chunks) {
    __shared__ Stage<FU, NUM> S;
    const unsigned w = xcd_remap(blockIdx.x, gridDim.x);
    const unsigned t = w % n_tiles, ch = w / n_tiles;
    const index_t per = (k + chunks - 1) / chunks;
    const index_t f0 = static_cast<index_t>(ch) * per;
    const index_t f1 = f0 + per < k ? f0 + per : k;
    if (f0 >= f1) return;  // whole workgroup
    tile_apply<E, FU, NUM, XT, YT>(A, X, Y, P, t, f0, f1, S);
}

/// A staged field value at byte offset o, in the accumulation type.
template <class XT, class V>
__device__ inline V lds_val(const char *base, unsigned o) {
    if constexpr (std::is_same_v<XT, std::uint16_t>)
        return static_cast<V>(__uint_as_float(static_cast<unsigned>(*reinterpret_cast<const std::uint16_t *>(base + o)) << 16));
    else
        return static_cast<V>(*reinterpret_cast<const XT *>(base + o));
}

/// Orders this wave's LDS accesses in program order for the compiler (the
/// hardware executes one wave's LDS instructions in order, so a lane's read
/// after another lane's write in program order sees the write).
__device__ inline void wave_lds_order() { asm volatile("" ::: "memory"); }

/// Wave-tile apply: each wave owns 64 consecutive segments and runs all its
/// fields [f0, f1) without workgroup barriers.  Per step of FU fields the
/// wave's distinct 16-byte source chunks (one or two per lane) go from
/// registers into the wave's own LDS area, then every lane forms its
/// segment's sums from LDS; the chunks of the fields kDepth steps ahead are
/// loaded into registers meanwhile.  Padding terms (past a segment's length,
/// up to the wave's longest) read a zero slot with zero weights.  The sums
/// of every row are those of the other apply kernels (CSR order per segment,
/// long-row partials reduced in segment order), so results are identical.
template <int E, int FU, int NUM, class XT, class YT>
__global__ void __launch_bounds__(256) spmm_wave_kernel(Mat A, const XT *X, YT *Y, double *P, index_t k, unsigned n_wt, int chunks) {
    using V = typename NumTraits<NUM>::V;
    using W = typename NumTraits<NUM>::W;
    constexpr int kDepth = FU >= 4 ? 1 : 4 / FU;
    constexpr index_t kAhead = static_cast<index_t>(kDepth) * FU;
    // field u's staged chunks at u * kStride, followed by 16 zero bytes (the
    // slot padding terms read)
    constexpr int kStride = kWaveStage + 16;
    constexpr unsigned kZero = kWaveStage;
    __shared__ __align__(16) char lds[4][FU * kStride];
    const int lane = static_cast<int>(threadIdx.x) & 63, wv = static_cast<int>(threadIdx.x) >> 6;
    const unsigned wid = xcd_remap(blockIdx.x, gridDim.x) * 4u + static_cast<unsigned>(wv);
    // chunk-major: the waves running at one time share their fields, whose
    // source lines then come from L2 (tile-major order measured slower)
    const unsigned t = wid % n_wt, ch = wid / n_wt;
    if (ch >= static_cast<unsigned>(chunks)) return;  // whole wave (no workgroup barriers in this kernel)
    const index_t per = (k + chunks - 1) / chunks;
    const index_t f0 = static_cast<index_t>(ch) * per;
    const index_t f1 = f0 + per < k ? f0 + per : k;
    if (f0 >= f1) return;
    const index_t i = static_cast<index_t>(t) * kWaveSeg + lane;
    const bool valid = i < A.n_seg;
    const int nu = A.wnu[t];
    if (nu < 0) {  // too many distinct chunks: direct gathers
        if (valid) segment_apply<E, 1, NUM>(A, X, Y, P, i, f0, f1);
        return;
    }
    index_t r = 0, b = 0;
    int n = 0, slot = -1;
    if (valid) {
        r = A.seg_row[i];
        b = A.seg_begin[i];
        n = A.seg_len[i];
        slot = A.seg_slot[i];
    }
    int n_max = n;
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) n_max = max(n_max, __shfl_xor(n_max, d, 64));
    const int n_wave = __builtin_amdgcn_readfirstlane(n_max);
    char *stg = lds[wv];
    if (lane < 4 * FU) *reinterpret_cast<unsigned *>(stg + (lane >> 2) * kStride + kZero + 4 * (lane & 3)) = 0u;
    unsigned lc[E];
    W h[E], l[E];
#pragma unroll
    for (int e = 0; e < E; ++e) {
        lc[e] = kZero;
        h[e] = W(0);
        l[e] = W(0);
        if (e < n) {
            lc[e] = static_cast<unsigned>(A.wlpos[b + e]) * static_cast<unsigned>(sizeof(XT));
            load_w<NUM>(A, b + e, h[e], l[e]);
        }
    }
    // this lane's chunks (lanes without one load the field's first chunk, whose
    // slot no nonzero reads)
    const std::int32_t *wc = A.wcol + static_cast<index_t>(t) * kWaveChunks;
    const bool two = nu > 64;  // wave-uniform
    const XT *xa = X + (lane < nu ? wc[lane] : 0);
    const XT *xb = X + (two && lane + 64 < nu ? wc[lane + 64] : 0);
    using u32x4 = unsigned __attribute__((ext_vector_type(4)));
    u32x4 va[kDepth][FU], vb[kDepth][FU];
#pragma unroll
    for (int d = 0; d < kDepth; ++d)
#pragma unroll
        for (int u = 0; u < FU; ++u) {
            const index_t fu = min(f0 + d * FU + u, f1 - 1);
            va[d][u] = *reinterpret_cast<const u32x4 *>(xa + fu * A.n_src);
            vb[d][u] = two ? *reinterpret_cast<const u32x4 *>(xb + fu * A.n_src) : u32x4{0u, 0u, 0u, 0u};
        }
    for (index_t fb = f0; fb < f1; fb += kAhead) {
#pragma unroll
        for (int d = 0; d < kDepth; ++d) {
            const index_t f = fb + d * FU;
            if (f < f1) {  // wave-uniform
                wave_lds_order();
#pragma unroll
                for (int u = 0; u < FU; ++u) {
                    *reinterpret_cast<u32x4 *>(stg + u * kStride + 16 * lane) = va[d][u];
                    if (two) *reinterpret_cast<u32x4 *>(stg + u * kStride + 16 * (lane + 64)) = vb[d][u];
                }
                wave_lds_order();
                if (f + kAhead < f1) {
#pragma unroll
                    for (int u = 0; u < FU; ++u) {
                        const index_t fu = min(f + kAhead + u, f1 - 1);
                        va[d][u] = *reinterpret_cast<const u32x4 *>(xa + fu * A.n_src);
                        if (two) vb[d][u] = *reinterpret_cast<const u32x4 *>(xb + fu * A.n_src);
                    }
                }
                RowAcc<NUM> acc[FU];
#pragma unroll
                for (int u = 0; u < FU; ++u) begin_row(acc[u], A, h, n);
#pragma unroll
                for (int g = 0; g < E; g += 4) {
                    if (g < n_wave) {
                        V xv[FU][4];
#pragma unroll
                        for (int u = 0; u < FU; ++u)
#pragma unroll
                            for (int q = 0; q < 4; ++q) xv[u][q] = lds_val<XT, V>(stg + u * kStride, lc[g + q]);
#pragma unroll
                        for (int u = 0; u < FU; ++u)
#pragma unroll
                            for (int q = 0; q < 4; ++q) acc[u].add(h[g + q], l[g + q], xv[u][q]);
                    }
                }
                if (valid) {
#pragma unroll
                    for (int u = 0; u < FU; ++u) {
                        if (f + u < f1) {
                            const double val = acc[u].value();
                            if (slot < 0)
                                st(Y + (f + u) * A.n_dst + r, val);
                            else
                                P[(f + u) * A.n_slots + slot] = val;
                        }
                    }
                }
            }
        }
    }
}

/// Y of the long rows: the partials of each row summed in segment order.
template <class YT>
__device__ inline void long_reduce(const Mat &A, const double *P, YT *Y, index_t k, index_t t) {
    const int j = static_cast<int>(t % A.n_long);
    const index_t f = t / A.n_long;
    double s = 0.0;
    for (int q = A.long_slot[j]; q < A.long_slot[j + 1]; ++q) s += P[f * A.n_slots + q];
    st(Y + f * A.n_dst + A.long_row[j], s);
}

template <class YT>
__global__ void long_reduce_kernel(Mat A, const double *P, YT *Y, index_t k) {
    const index_t t = static_cast<index_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= static_cast<index_t>(A.n_long) * k) return;
    long_reduce(A, P, Y, k, t);
}

/// Long rows of the special modes: one lane per (long row, field), the whole
/// row in CSR order (loads issued eight entries ahead of the sums).
template <int NUM>
__global__ void long_rows_special_kernel(Mat A, const index_t *row_ptr, const double *X, double *Y, index_t k) {
    const index_t t = static_cast<index_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= static_cast<index_t>(A.n_long) * k) return;
    const index_t j = t % A.n_long, f = t / A.n_long;
    const index_t r = A.long_row[j];
    if (!special_writes(A, r, -1)) return;
    const index_t b = row_ptr[r], e = row_ptr[r + 1];
    const double *xf = X + f * A.n_src;
    RowAcc<NUM> acc;
    if constexpr (NUM == kNanAware) {
        double tot = 0.0;
        for (index_t q = b; q < e; ++q) tot += A.val[q];
        acc.total = tot;
        acc.thres = A.thres;
    }
    index_t q = b;
    for (; q + 8 <= e; q += 8) {
        double w[8], x[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            w[u] = A.val[q + u];
            x[u] = xf[A.col[q + u]];
        }
#pragma unroll
        for (int u = 0; u < 8; ++u) acc.add(w[u], 0.0, x[u]);
    }
    for (; q < e; ++q) acc.add(A.val[q], 0.0, xf[A.col[q]]);
    Y[f * A.n_dst + r] = acc.value();
}

inline Mat mat_of(const InterpolationMatrix &A) {
    Mat M{};
    M.col = A.col32.get();
    M.val = A.val.get();
    M.hi = A.val_hi.get();
    M.lo = A.val_lo.get();
    M.n_src = A.n_src;
    M.n_dst = A.n_dst;
    M.seg_row = A.seg_row.get();
    M.seg_begin = A.seg_begin.get();
    M.seg_len = A.seg_len.get();
    M.seg_slot = A.seg_slot.get();
    M.n_seg = A.n_seg;
    M.long_row = A.long_rows.get();
    M.long_slot = A.long_slot.get();
    M.n_long = A.n_long_rows;
    M.n_slots = A.n_slots;
    M.tile_nu = A.tile_nu.get();
    M.tile_col = A.tile_col.get();
    M.lcol = A.t_lcol.get();
    M.n_tiles = A.n_tiles;
    M.thres = 0.0;
    M.mask = nullptr;
    M.wnu = nullptr;
    M.wcol = nullptr;
    M.wlpos = nullptr;
    M.n_wt = A.n_wtiles;
    return M;
}

/// Whether the wave-tile kernel can serve fields of XT: chunks are read with
/// 16-byte loads, so the fields must be 16-byte aligned and hold whole chunks.
template <class XT>
inline bool wave_ok(const InterpolationMatrix &A, const void *x) {
    return A.n_src > 0 && A.n_src % chunk_cols<XT>() == 0 && (reinterpret_cast<std::uintptr_t>(x) % 16) == 0;
}

/// Whether work enqueued on s is being recorded into a graph: nothing may be
/// allocated, synchronised or memset then.
inline bool capturing(hipStream_t s) {
    hipStreamCaptureStatus st = hipStreamCaptureStatusNone;
    AXIS_HIP_CHECK(hipStreamIsCapturing(s, &st));
    return st != hipStreamCaptureStatusNone;
}

/// FP64 partial-sum workspace for k fields of applies on stream s, owned by
/// the matrix (one per stream; the caller holds A.apply_mu).  A stream under
/// graph capture must find it allocated by an earlier apply on that stream.
inline double *partials(InterpolationMatrix &A, index_t k, hipStream_t s) {
    if (A.n_slots == 0) return nullptr;
    const std::size_t need = static_cast<std::size_t>(A.n_slots) * static_cast<std::size_t>(k);
    DeviceBuffer<double> &ws = A.partial_ws[s];
    if (ws.size() < need) {
        if (capturing(s))
            throw InvalidArgument("apply: a stream under graph capture needs its long-row workspace from an earlier apply of as many fields on it");
        ws.resize(need);
    }
    return ws.get();
}

/// Extra operands of the special modes.
struct SpecialArgs {
    double thres = 0.0;
    const std::int32_t *mask = nullptr;
    double *partials = nullptr;  // long-row partial sums (null: the matrix's own workspace)
};

template <int E, int FU, int NUM, class XT, class YT>
void launch_e(InterpolationMatrix &A, const void *x, void *y, index_t k, int chunks, bool tiled, bool wave, hipStream_t s, const SpecialArgs &sa) {
    Mat M = mat_of(A);
    M.thres = sa.thres;
    M.mask = sa.mask;
    double *P = sa.partials ? sa.partials : partials(A, k, s);
    const unsigned nb = static_cast<unsigned>(std::max<index_t>(tiled ? A.n_tiles : (A.n_seg + kBlock - 1) / kBlock, 1));
    const std::uint64_t nwork = static_cast<std::uint64_t>(nb) * chunks;
    if (nwork >= (1ull << 31)) throw InvalidArgument("apply: launch too large");
    if (wave && A.n_seg > 0) {
        const int ws = width_slot(chunk_cols<XT>());
        if (!A.wave_ready[ws]) throw InvalidArgument("apply: wave tiles not built");
        M.wnu = A.wnu[ws].get();
        M.wcol = A.wcol[ws].get();
        M.wlpos = A.wlpos[ws].get();
        const std::uint64_t items = static_cast<std::uint64_t>(A.n_wtiles) * chunks;
        if (items >= (1ull << 33)) throw InvalidArgument("apply: launch too large");
        spmm_wave_kernel<E, FU, NUM, XT, YT><<<static_cast<unsigned>((items + 3) / 4), 256, 0, s>>>(
            M, static_cast<const XT *>(x), static_cast<YT *>(y), P, k, static_cast<unsigned>(A.n_wtiles), chunks);
        AXIS_LAUNCH_CHECK();
    } else if (A.n_seg > 0) {
        if (tiled)
            spmm_tile_kernel<E, FU, NUM, XT, YT><<<static_cast<unsigned>(nwork), kTileSeg, 0, s>>>(M, static_cast<const XT *>(x),
                                                                                                    static_cast<YT *>(y), P, k, nb, chunks);
        else
            spmm_seg_kernel<E, FU, NUM, XT, YT><<<static_cast<unsigned>(nwork), kBlock, 0, s>>>(M, static_cast<const XT *>(x),
                                                                                                 static_cast<YT *>(y), P, k, nb, chunks);
        AXIS_LAUNCH_CHECK();
    }
    if (A.n_long_rows > 0) {
        const index_t n = static_cast<index_t>(A.n_long_rows) * k;
        if constexpr (is_special(NUM))
            long_rows_special_kernel<NUM><<<grid_for(n, 256), 256, 0, s>>>(M, A.row_ptr.get(), static_cast<const double *>(x),
                                                                          static_cast<double *>(y), k);
        else
            long_reduce_kernel<YT><<<grid_for(n, 256), 256, 0, s>>>(M, P, static_cast<YT *>(y), k);
        AXIS_LAUNCH_CHECK();
    }
}

template <int FU, int NUM, class XT, class YT>
void launch_fu(InterpolationMatrix &A, const void *x, void *y, index_t k, int chunks, bool tiled, bool wave, hipStream_t s, const SpecialArgs &sa) {
    switch (A.seg_cap) {
        case 4: return launch_e<4, FU, NUM, XT, YT>(A, x, y, k, chunks, tiled, wave, s, sa);
        case 8: return launch_e<8, FU, NUM, XT, YT>(A, x, y, k, chunks, tiled, wave, s, sa);
        case 12: return launch_e<12, FU, NUM, XT, YT>(A, x, y, k, chunks, tiled, wave, s, sa);
        default: return launch_e<16, FU, NUM, XT, YT>(A, x, y, k, chunks, tiled, wave, s, sa);
    }
}

template <int NUM, class XT, class YT>
void launch_xy(InterpolationMatrix &A, const void *x, void *y, index_t k, const Launch &L, hipStream_t s, const SpecialArgs &sa = SpecialArgs{}) {
    const bool wave = L.wave && !is_special(NUM) && wave_ok<XT>(A, x);
    const bool tiled = L.tiled || (L.wave && !wave);  // a wave shape on fields it cannot serve runs staged tiles
    switch (L.fu) {
        case 1: return launch_fu<1, NUM, XT, YT>(A, x, y, k, L.chunks, tiled, wave, s, sa);
        case 2: return launch_fu<2, NUM, XT, YT>(A, x, y, k, L.chunks, tiled, wave, s, sa);
        case 4: return launch_fu<4, NUM, XT, YT>(A, x, y, k, L.chunks, tiled, wave, s, sa);
        default: throw InvalidArgument("apply: unsupported fields per step");
    }
}

template <int NUM>
void launch_num(InterpolationMatrix &A, const void *x, FieldType xt, void *y, FieldType yt, index_t k, const Launch &L, hipStream_t s,
                const SpecialArgs &sa) {
    if (xt == FieldType::F32 && yt == FieldType::F32) return launch_xy<NUM, float, float>(A, x, y, k, L, s, sa);
    if (xt == FieldType::F32 && yt == FieldType::F64) return launch_xy<NUM, float, double>(A, x, y, k, L, s, sa);
    if (xt == FieldType::BF16 && yt == FieldType::F32) return launch_xy<NUM, std::uint16_t, float>(A, x, y, k, L, s, sa);
    if (xt == FieldType::BF16 && yt == FieldType::F64) return launch_xy<NUM, std::uint16_t, double>(A, x, y, k, L, s, sa);
    throw InvalidArgument("apply: unsupported field type combination");
}

/// One apply with launch shape L; sa.partials, when given, replaces the
/// matrix's own long-row workspace (n_slots * k doubles).
inline int chunk_cols_of(FieldType xt) { return xt == FieldType::F64 ? 2 : (xt == FieldType::F32 ? 4 : 8); }

/// Whether launch shape L on fields x of type xt runs wave tiles (a wave
/// shape on fields they cannot serve runs staged tiles instead).
inline bool runs_wave(const InterpolationMatrix &A, const Launch &L, FieldType xt, const void *x) {
    const int cw = chunk_cols_of(xt);
    return L.wave && A.n_src > 0 && A.n_src % cw == 0 && (reinterpret_cast<std::uintptr_t>(x) % 16) == 0;
}

/// Whether the layouts launch shape L needs on these fields exist already.
inline bool layout_ready(const InterpolationMatrix &A, const Launch &L, FieldType xt, const void *x) {
    if (!A.apply_ready) return false;
    if (runs_wave(A, L, xt, x)) return A.wave_ready[width_slot(chunk_cols_of(xt))];
    return !(L.tiled || L.wave) || A.tiles_ready;
}

inline void launch(InterpolationMatrix &A, const void *x, FieldType xt, void *y, FieldType yt, index_t k, ApplyMode mode, const Launch &L,
                   hipStream_t s, const SpecialArgs &sa = SpecialArgs{}) {
    const bool wave_fit = runs_wave(A, L, xt, x);
    if (L.tiled || (L.wave && !wave_fit)) A.prepare_tiles(s);
    if (wave_fit) A.prepare_wave_tiles(chunk_cols_of(xt), s);
    if (mode == ApplyMode::Plain) return launch_xy<kPlain, double, double>(A, x, y, k, L, s, sa);
    if (mode == ApplyMode::Mixed) return launch_num<kComp32>(A, x, xt, y, yt, k, L, s, sa);
    return launch_num<kAcc64>(A, x, xt, y, yt, k, L, s, sa);
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

/// Without a tuned entry: enough field chunks to give every CU a few
/// workgroups, 4 fields per step (2 for FP64, whose gathers are twice as wide).
inline Launch heuristic(const InterpolationMatrix &A, index_t k, ApplyMode mode) {
    Launch L;
    L.tiled = k >= 8;
    L.fu = (mode == ApplyMode::Mixed) ? 4 : 2;
    if (k < L.fu) L.fu = k >= 2 ? 2 : 1;
    const index_t nb = (A.n_seg + kBlock - 1) / kBlock;
    index_t want = (2048 + nb - 1) / std::max<index_t>(nb, 1);  // ~8 workgroups per CU
    L.chunks = static_cast<int>(std::max<index_t>(1, std::min<index_t>(want, (k + L.fu - 1) / L.fu)));
    return L;
}

inline Launch choose(const InterpolationMatrix &A, index_t k, FieldType xt, FieldType yt, ApplyMode mode) {
    std::lock_guard<std::mutex> g(tune_mutex());
    auto it = tune_cache().find(TuneKey{A.n_dst, A.nnz, k, static_cast<int>(xt), static_cast<int>(yt), static_cast<int>(mode)});
    if (it != tune_cache().end()) return it->second;
    return heuristic(A, k, mode);
}

}  // namespace axis_hip::apply_detail
