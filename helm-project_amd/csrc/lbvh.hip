// lbvh.hip -- GPU linear BVH broad phase over float cell boxes.
//
// Replaces the reference's ArborX BoundingVolumeHierarchy build + intersects
// query (libs/axis/src/solver/weight_generator.cpp:3345-3367).  Result: for
// every destination cell the list of source cells whose float boxes
// intersect its box (closed intervals, as ArborX::intersects).
//
//   1. 30-bit Morton codes of the box centres on unit-sphere xyz
//      (10 bits per axis over [-1.1, 1.1]^3),
//   2. LDS radix sort of (code, cell) pairs,
//   3. Karras radix-tree build (Karras 2012, "Maximizing parallelism in the
//      construction of BVHs, octrees and k-d trees"), duplicate codes broken
//      by sorted position,
//   4. bottom-up box refit with one atomic arrival counter per internal node,
//   5. one stack traversal per query that keeps up to 32 hits; the queries
//      with more (rows around the poles of a lat-lon source meet thousands)
//      are counted and filled by one wave (or workgroup) each,
//      breadth-first; nodes are packed with both child boxes.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "axis_hip.hpp"
#include "host_util.hpp"
#include "morton.hpp"

namespace axis_hip::detail {

namespace {

constexpr int kBlock = 256;
constexpr int kWave = 64;
constexpr int kStack = 64;
constexpr std::uint32_t kLeaf = 0x80000000u;

struct FBox {
    float lo[3], hi[3];
};

__device__ inline FBox load_box(const float *b, std::int64_t i) {
    FBox x;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        x.lo[a] = b[6 * i + a];
        x.hi[a] = b[6 * i + 3 + a];
    }
    return x;
}

__device__ inline bool overlap(const FBox &a, const FBox &b) {
    return !(a.hi[0] < b.lo[0] || b.hi[0] < a.lo[0] || a.hi[1] < b.lo[1] || b.hi[1] < a.lo[1] || a.hi[2] < b.lo[2] || b.hi[2] < a.lo[2]);
}

__global__ void morton_kernel(const float *boxes, std::int64_t n, std::uint32_t *codes, std::uint32_t *ids) {
    std::int64_t i = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    FBox b = load_box(boxes, i);
    codes[i] = morton30(0.5f * (b.lo[0] + b.hi[0]), 0.5f * (b.lo[1] + b.hi[1]), 0.5f * (b.lo[2] + b.hi[2]));
    ids[i] = static_cast<std::uint32_t>(i);
}

// common prefix length of sorted keys i and j, ties broken by position
__device__ inline int delta(const std::uint32_t *codes, int n, int i, int j) {
    if (j < 0 || j >= n) return -1;
    std::uint32_t a = codes[i], b = codes[j];
    if (a == b) return 32 + __clz(static_cast<std::uint32_t>(i ^ j));
    return __clz(a ^ b);
}

// internal node i (0..n-2): children and parents
__global__ void karras_kernel(const std::uint32_t *codes, int n, std::uint32_t *left, std::uint32_t *right, int *parent, int *rlo, int *rhi) {
    int i = static_cast<int>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n - 1) return;
    int d = (delta(codes, n, i, i + 1) - delta(codes, n, i, i - 1)) >= 0 ? 1 : -1;
    int dmin = delta(codes, n, i, i - d);
    int lmax = 2;
    while (delta(codes, n, i, i + lmax * d) > dmin) lmax <<= 1;
    int l = 0;
    for (int t = lmax >> 1; t >= 1; t >>= 1)
        if (delta(codes, n, i, i + (l + t) * d) > dmin) l += t;
    int j = i + l * d;
    int dnode = delta(codes, n, i, j);
    int s = 0;
    int t = l;
    do {
        t = (t + 1) >> 1;
        if (delta(codes, n, i, i + (s + t) * d) > dnode) s += t;
    } while (t > 1);
    int gamma = i + s * d + (d < 0 ? d : 0);
    int lo = i < j ? i : j, hi = i < j ? j : i;
    rlo[i] = lo;
    rhi[i] = hi;
    std::uint32_t lc = (lo == gamma) ? (kLeaf | static_cast<std::uint32_t>(gamma)) : static_cast<std::uint32_t>(gamma);
    std::uint32_t rc = (hi == gamma + 1) ? (kLeaf | static_cast<std::uint32_t>(gamma + 1)) : static_cast<std::uint32_t>(gamma + 1);
    left[i] = lc;
    right[i] = rc;
    // parents: internal nodes at [0, n-1), leaves at [n-1, 2n-1)
    parent[(lc & kLeaf) ? (n - 1 + (lc & ~kLeaf)) : lc] = i;
    parent[(rc & kLeaf) ? (n - 1 + (rc & ~kLeaf)) : rc] = i;
}

// Bottom-up box refit in two phases, so that no arrival counter is shared
// between workgroups (an agent-scope fence per tree level costs an L2
// write-back on this GPU):
//   1. each workgroup owns a chunk of kRefitChunk sorted leaves and climbs
//      from its leaves through the internal nodes whose leaf range lies inside
//      the chunk, with workgroup-scope arrival counters (the second child to
//      arrive merges the two child boxes and climbs on);
//   2. one workgroup finishes the nodes whose range crosses a chunk boundary,
//      starting from their children that phase 1 completed.
constexpr int kRefitChunk = 4096;
constexpr int kRefitBlock = 1024;

__device__ inline void merge_node(const float *boxes, const std::uint32_t *ids, const std::uint32_t *left, const std::uint32_t *right,
                                  float *node_box, int p) {
    FBox b;
    std::uint32_t c[2] = {left[p], right[p]};
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        b.lo[a] = INFINITY;
        b.hi[a] = -INFINITY;
    }
    for (int h = 0; h < 2; ++h) {
        FBox x = (c[h] & kLeaf) ? load_box(boxes, ids[c[h] & ~kLeaf]) : load_box(node_box, c[h]);
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            b.lo[a] = fminf(b.lo[a], x.lo[a]);
            b.hi[a] = fmaxf(b.hi[a], x.hi[a]);
        }
    }
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        node_box[6 * static_cast<std::int64_t>(p) + a] = b.lo[a];
        node_box[6 * static_cast<std::int64_t>(p) + 3 + a] = b.hi[a];
    }
}

__global__ void __launch_bounds__(kRefitBlock) refit_chunk_kernel(const float *boxes, const std::uint32_t *ids, int n, const std::uint32_t *left,
                                                             const std::uint32_t *right, const int *parent, const int *rlo, const int *rhi,
                                                             float *node_box, int *arrive) {
    const int c0 = static_cast<int>(blockIdx.x) * kRefitChunk;
    const int c1 = min(c0 + kRefitChunk, n);
    for (int k = c0 + static_cast<int>(threadIdx.x); k < c1; k += kRefitBlock) {
        int p = parent[n - 1 + k];
        while (p >= 0 && rlo[p] >= c0 && rhi[p] < c1) {
            if (__hip_atomic_fetch_add(&arrive[p], 1, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_WORKGROUP) == 0) break;
            merge_node(boxes, ids, left, right, node_box, p);
            p = parent[p];
        }
    }
}

__device__ inline bool crosses_chunk(const int *rlo, const int *rhi, int p) { return rlo[p] / kRefitChunk != rhi[p] / kRefitChunk; }

// children of chunk-crossing nodes that phase 1 completed (leaves included)
__global__ void refit_frontier_kernel(int n, const std::uint32_t *left, const std::uint32_t *right, const int *rlo, const int *rhi, int *list,
                                      int *count) {
    const int p = static_cast<int>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (p >= n - 1 || !crosses_chunk(rlo, rhi, p)) return;
    const std::uint32_t c[2] = {left[p], right[p]};
    for (int h = 0; h < 2; ++h) {
        const bool leaf = (c[h] & kLeaf) != 0;
        if (leaf || !crosses_chunk(rlo, rhi, static_cast<int>(c[h]))) {
            // encode leaves as n - 1 + leaf index, internal nodes as themselves
            list[atomicAdd(count, 1)] = leaf ? n - 1 + static_cast<int>(c[h] & ~kLeaf) : static_cast<int>(c[h]);
        }
    }
}

constexpr int kTopBlock = 1024;

__global__ void __launch_bounds__(kTopBlock) refit_top_kernel(const float *boxes, const std::uint32_t *ids, int n, const std::uint32_t *left,
                                                              const std::uint32_t *right, const int *parent, float *node_box, int *arrive,
                                                              const int *list, const int *count) {
    const int m = *count;
    for (int t = static_cast<int>(threadIdx.x); t < m; t += kTopBlock) {
        int p = parent[list[t]];
        while (p >= 0) {
            if (__hip_atomic_fetch_add(&arrive[p], 1, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_WORKGROUP) == 0) break;
            merge_node(boxes, ids, left, right, node_box, p);
            p = parent[p];
        }
    }
}

// Traversal layout: one 64-byte record per internal node holding both child
// boxes and child references (internal node index, or kLeaf | source cell),
// so that visiting a node is one dependent load and testing a leaf none.
struct __align__(16) Node {
    float box[2][6];  // per child: lo xyz, hi xyz
    std::uint32_t child[2];
    std::uint32_t pad[2];
};

__global__ void pack_kernel(const float *boxes, const std::uint32_t *ids, int n, const std::uint32_t *left, const std::uint32_t *right,
                            const float *node_box, Node *nodes) {
    const int p = static_cast<int>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (p >= n - 1) return;
    Node x;
    const std::uint32_t c[2] = {left[p], right[p]};
    for (int h = 0; h < 2; ++h) {
        const bool leaf = (c[h] & kLeaf) != 0;
        const std::uint32_t id = leaf ? ids[c[h] & ~kLeaf] : c[h];
        const FBox b = leaf ? load_box(boxes, id) : load_box(node_box, c[h]);
        for (int a = 0; a < 3; ++a) {
            x.box[h][a] = b.lo[a];
            x.box[h][3 + a] = b.hi[a];
        }
        x.child[h] = leaf ? (kLeaf | id) : id;
    }
    x.pad[0] = x.pad[1] = 0;
    nodes[p] = x;
}

__device__ inline bool overlap_child(const FBox &q, const Node &x, int h) {
    const float *b = x.box[h];
    return !(q.hi[0] < b[0] || b[3] < q.lo[0] || q.hi[1] < b[1] || b[4] < q.lo[1] || q.hi[2] < b[2] || b[5] < q.lo[2]);
}

// Per-lane stack traversal of query j (rows[qi] when a row list is given).
//   kMode 0: count all hits;  kMode 1: write all hits at offsets[j];
//   kMode 2: first pass over every query -- keep up to kCap hits in
//            tmp[j * kCap ...]; a query that meets more stops there, is
//            appended to `over` and finished by the workgroup kernel below.
// Stack overflow (never expected: depth <= 62 with 30-bit codes + position
// tie-break) sets *err so the host reports it instead of returning a short list.
constexpr int kCap = 32;

template <int kMode>
__global__ void __launch_bounds__(kBlock) query_kernel(const Node *nodes, const float *src_boxes, int n_src, const float *dst_boxes, std::int64_t n_q,
                                                       const std::int32_t *rows, std::int64_t *count, const std::int64_t *offsets, std::int32_t *out,
                                                       std::int32_t *over, int *n_over, int *err) {
    const std::int64_t qi = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (qi >= n_q) return;
    const std::int64_t j = rows ? rows[qi] : qi;
    const FBox q = load_box(dst_boxes, j);
    std::int64_t cnt = 0;
    const std::int64_t w = kMode == 1 ? offsets[j] : (kMode == 2 ? j * kCap : 0);
    bool full = false;
    if (n_src == 1) {
        if (overlap(q, load_box(src_boxes, 0))) {
            if (kMode != 0) out[w] = 0;
            ++cnt;
        }
    } else if (n_src > 1) {
        // descend into an overlapping child directly, stacking the other one
        // when both overlap; pop only at dead ends
        std::uint32_t stack[kStack];
        int sp = 0;
        std::uint32_t nd = 0;
        while (true) {
            const Node x = nodes[nd];
            bool inner[2] = {false, false};
            for (int h = 0; h < 2; ++h) {
                if (!overlap_child(q, x, h)) continue;
                const std::uint32_t c = x.child[h];
                if (!(c & kLeaf)) {
                    inner[h] = true;
                    continue;
                }
                if (kMode == 2 && cnt == kCap) {
                    full = true;
                    break;
                }
                if (kMode != 0) out[w + cnt] = static_cast<std::int32_t>(c & ~kLeaf);
                ++cnt;
            }
            if (full) break;
            if (inner[0] && inner[1]) {
                if (sp < kStack)
                    stack[sp++] = x.child[1];
                else
                    atomicOr(err, 1);
                nd = x.child[0];
            } else if (inner[0]) {
                nd = x.child[0];
            } else if (inner[1]) {
                nd = x.child[1];
            } else {
                if (sp == 0) break;
                nd = stack[--sp];
            }
        }
    }
    if (kMode == 2 && full) {
        count[j] = kCap + 1;  // replaced by the exact count before the scan
        over[atomicAdd(n_over, 1)] = static_cast<std::int32_t>(j);
        return;
    }
    if (kMode != 1) count[j] = cnt;
}

// Queries with more than kCap hits: one workgroup of kThreads per query,
// breadth-first over the tree, each thread taking one frontier node per step;
// hits and next-level nodes are placed by ballot prefix counts within each
// wave plus the counts of the lower waves, so the hit order is fixed and the
// count and fill passes agree.  Rows start at level 0 (one wave, a 2048-node
// frontier); a level wider than the frontier moves a row up one level (it is
// counted and filled there instead): level 1 is a 256-thread workgroup with a
// 6144-node frontier, level 2 the per-lane kernel.
template <bool kFill, int kThreads, int kFrontCap, int kLevel>
__global__ void __launch_bounds__(kThreads) query_long_kernel(const Node *nodes, const float *dst_boxes, const std::int32_t *rows,
                                                              std::int64_t *count, const std::int64_t *offsets, std::int32_t *out,
                                                              std::uint8_t *level, std::int32_t *next_rows, int *n_next) {
    constexpr int kWaves = kThreads / kWave;
    __shared__ std::uint32_t front[2][kFrontCap];
    __shared__ int wsum[2][kWaves];
    const int tid = static_cast<int>(threadIdx.x), lane = tid & (kWave - 1), wv = tid / kWave;
    const std::int64_t j = rows[blockIdx.x];
    if (kFill && level[j] != kLevel) return;
    const FBox q = load_box(dst_boxes, j);
    const std::uint64_t below = (std::uint64_t(1) << lane) - 1;
    const std::int64_t w = kFill ? offsets[j] : 0;
    std::int64_t cnt = 0;
    if (tid == 0) front[0][0] = 0;
    __syncthreads();
    int cur = 0, m = 1;
    bool wide = false;
    while (m > 0) {
        int nm = 0;
        for (int i0 = 0; i0 < m; i0 += kThreads) {
            const bool act = i0 + tid < m;
            Node x;
            if (act) x = nodes[front[cur][i0 + tid]];
            for (int h = 0; h < 2; ++h) {
                const bool ov = act && overlap_child(q, x, h);
                const std::uint32_t c = act ? x.child[h] : 0u;
                const bool hit = ov && (c & kLeaf), push = ov && !(c & kLeaf);
                const std::uint64_t mh = __ballot(hit), mp = __ballot(push);
                int hb = 0, pb = 0, ht = __popcll(mh), pt = __popcll(mp);
                if (kWaves > 1) {
                    if (lane == 0) {
                        wsum[0][wv] = ht;
                        wsum[1][wv] = pt;
                    }
                    __syncthreads();
                    ht = pt = 0;
                    for (int v = 0; v < kWaves; ++v) {
                        if (v < wv) {
                            hb += wsum[0][v];
                            pb += wsum[1][v];
                        }
                        ht += wsum[0][v];
                        pt += wsum[1][v];
                    }
                }
                if (kFill && hit) out[w + cnt + hb + __popcll(mh & below)] = static_cast<std::int32_t>(c & ~kLeaf);
                const int at = nm + pb + __popcll(mp & below);
                if (push && at < kFrontCap) front[cur ^ 1][at] = c;
                cnt += ht;
                nm += pt;
                __syncthreads();
            }
        }
        if (nm > kFrontCap) {
            wide = true;
            break;
        }
        cur ^= 1;
        m = nm;
    }
    if (kFill || tid != 0) return;
    if (wide) {
        level[j] = kLevel + 1;
        next_rows[atomicAdd(n_next, 1)] = static_cast<std::int32_t>(j);
    } else {
        count[j] = cnt;
    }
}

constexpr int kFront0 = 2048, kFront1 = 6144, kBlock1 = 256;

// first-pass hits of the queries with at most kCap of them -> candidate CSR
// (kCap consecutive lanes per query, so reads and writes are contiguous)
__global__ void compact_hits_kernel(const std::int32_t *tmp, const std::int64_t *count, const std::int64_t *offsets, std::int64_t n,
                                    std::int32_t *out) {
    const std::int64_t g = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    const std::int64_t j = g / kCap;
    const int e = static_cast<int>(g % kCap);
    if (j >= n) return;
    const std::int64_t c = count[j];
    if (c > kCap || e >= c) return;
    out[offsets[j] + e] = tmp[j * kCap + e];
}

}  // namespace

void lbvh_build(const float *src_boxes, index_t n_src, LbvhTree &t, hipStream_t s) {
    if (n_src >= (index_t(1) << 31) - 1) throw InvalidArgument("LBVH: more than 2^31 source cells");
    const int n = static_cast<int>(n_src);
    PhaseTrace tr("lbvh", s);
    DeviceBuffer<std::uint32_t> codes(std::max(n, 1)), ids(std::max(n, 1));
    DeviceBuffer<std::uint32_t> left(std::max(n - 1, 1)), right(std::max(n - 1, 1));
    DeviceBuffer<int> parent(std::max(2 * n - 1, 1)), arrive(std::max(n - 1, 1));
    DeviceBuffer<int> rlo(std::max(n - 1, 1)), rhi(std::max(n - 1, 1)), frontier(std::max(2 * n, 2)), n_frontier(1);
    DeviceBuffer<float> node_box(6 * std::max(n - 1, 1));
    if (n > 0) {
        morton_kernel<<<grid_for(n, kBlock), kBlock, 0, s>>>(src_boxes, n, codes.get(), ids.get());
        AXIS_LAUNCH_CHECK();
        radix_sort_pairs_u32(codes.get(), ids.get(), n, 30, s);
    }
    tr.mark("morton_sort");
    if (n > 1) {
        AXIS_HIP_CHECK(hipMemsetAsync(parent.get(), 0xFF, sizeof(int) * (2 * n - 1), s));  // root's parent = -1
        AXIS_HIP_CHECK(hipMemsetAsync(arrive.get(), 0, sizeof(int) * (n - 1), s));
        karras_kernel<<<grid_for(n - 1, kBlock), kBlock, 0, s>>>(codes.get(), n, left.get(), right.get(), parent.get(), rlo.get(), rhi.get());
        AXIS_LAUNCH_CHECK();
        refit_chunk_kernel<<<static_cast<unsigned>((n + kRefitChunk - 1) / kRefitChunk), kRefitBlock, 0, s>>>(
            src_boxes, ids.get(), n, left.get(), right.get(), parent.get(), rlo.get(), rhi.get(), node_box.get(), arrive.get());
        AXIS_LAUNCH_CHECK();
        AXIS_HIP_CHECK(hipMemsetAsync(n_frontier.get(), 0, sizeof(int), s));
        refit_frontier_kernel<<<grid_for(n - 1, kBlock), kBlock, 0, s>>>(n, left.get(), right.get(), rlo.get(), rhi.get(), frontier.get(),
                                                                         n_frontier.get());
        AXIS_LAUNCH_CHECK();
        refit_top_kernel<<<1, kTopBlock, 0, s>>>(src_boxes, ids.get(), n, left.get(), right.get(), parent.get(), node_box.get(), arrive.get(),
                                                 frontier.get(), n_frontier.get());
        AXIS_LAUNCH_CHECK();
    }
    t.n = n;
    t.src_boxes = src_boxes;
    t.nodes.resize(sizeof(Node) * static_cast<std::size_t>(std::max(n - 1, 1)));
    Node *nodes = reinterpret_cast<Node *>(t.nodes.get());
    if (n > 1) {
        pack_kernel<<<grid_for(n - 1, kBlock), kBlock, 0, s>>>(src_boxes, ids.get(), n, left.get(), right.get(), node_box.get(), nodes);
        AXIS_LAUNCH_CHECK();
    }
    tr.mark("build_refit");
}

void lbvh_query(const LbvhTree &t, const float *dst_boxes, index_t n_dst, DeviceBuffer<index_t> &cand_off, DeviceBuffer<std::int32_t> &cand_src,
                hipStream_t s) {
    const int n = t.n;
    const float *src_boxes = t.src_boxes;
    const Node *nodes = reinterpret_cast<const Node *>(t.nodes.get());
    PhaseTrace tr("lbvh", s);
    cand_off.resize(n_dst + 1);
    DeviceBuffer<int> err(1);
    AXIS_HIP_CHECK(hipMemsetAsync(err.get(), 0, sizeof(int), s));
    const index_t nd1 = std::max<index_t>(n_dst, 1);
    DeviceBuffer<index_t> count(nd1);
    DeviceBuffer<std::int32_t> tmp(static_cast<std::size_t>(nd1) * kCap), over(nd1), rows1(nd1), rows2(nd1);
    DeviceBuffer<std::uint8_t> level(nd1);
    DeviceBuffer<int> n_list(3);  // queries over kCap, moved to level 1, moved to level 2
    AXIS_HIP_CHECK(hipMemsetAsync(n_list.get(), 0, 3 * sizeof(int), s));
    if (n_dst > 0) {
        query_kernel<2><<<grid_for(n_dst, kBlock), kBlock, 0, s>>>(nodes, src_boxes, n, dst_boxes, n_dst, nullptr, count.get(), nullptr,
                                                                  tmp.get(), over.get(), n_list.get(), err.get());
        AXIS_LAUNCH_CHECK();
    }
    int h_list[3] = {0, 0, 0}, h_err = 0;
    AXIS_HIP_CHECK(hipMemcpyAsync(h_list, n_list.get(), sizeof h_list, hipMemcpyDeviceToHost, s));
    AXIS_HIP_CHECK(hipMemcpyAsync(&h_err, err.get(), sizeof(int), hipMemcpyDeviceToHost, s));
    AXIS_HIP_CHECK(hipStreamSynchronize(s));
    if (h_err) throw RuntimeError("LBVH traversal stack overflow");
    if (h_list[0] > 0) {
        AXIS_HIP_CHECK(hipMemsetAsync(level.get(), 0, static_cast<std::size_t>(nd1), s));
        query_long_kernel<false, kWave, kFront0, 0><<<static_cast<unsigned>(h_list[0]), kWave, 0, s>>>(
            nodes, dst_boxes, over.get(), count.get(), nullptr, nullptr, level.get(), rows1.get(), n_list.get() + 1);
        AXIS_LAUNCH_CHECK();
        AXIS_HIP_CHECK(hipMemcpyAsync(h_list, n_list.get(), sizeof h_list, hipMemcpyDeviceToHost, s));
        AXIS_HIP_CHECK(hipStreamSynchronize(s));
        if (h_list[1] > 0) {
            query_long_kernel<false, kBlock1, kFront1, 1><<<static_cast<unsigned>(h_list[1]), kBlock1, 0, s>>>(
                nodes, dst_boxes, rows1.get(), count.get(), nullptr, nullptr, level.get(), rows2.get(), n_list.get() + 2);
            AXIS_LAUNCH_CHECK();
            AXIS_HIP_CHECK(hipMemcpyAsync(h_list, n_list.get(), sizeof h_list, hipMemcpyDeviceToHost, s));
            AXIS_HIP_CHECK(hipStreamSynchronize(s));
            if (h_list[2] > 0) {
                query_kernel<0><<<grid_for(h_list[2], kBlock), kBlock, 0, s>>>(nodes, src_boxes, n, dst_boxes, h_list[2], rows2.get(),
                                                                              count.get(), nullptr, nullptr, nullptr, nullptr, err.get());
                AXIS_LAUNCH_CHECK();
            }
        }
    }
    exclusive_scan_i64(count.get(), cand_off.get(), n_dst, s);
    index_t total = 0;
    AXIS_HIP_CHECK(hipMemcpyAsync(&total, cand_off.get() + n_dst, sizeof(index_t), hipMemcpyDeviceToHost, s));
    AXIS_HIP_CHECK(hipMemcpyAsync(&h_err, err.get(), sizeof(int), hipMemcpyDeviceToHost, s));
    AXIS_HIP_CHECK(hipStreamSynchronize(s));
    tr.mark("query_count");
    if (h_err) throw RuntimeError("LBVH traversal stack overflow");
    // the reference stores candidate offsets as int (weight_generator.cpp:247-254)
    if (total >= (index_t(1) << 31)) throw RuntimeError("LBVH: candidate pair count exceeds 2^31");
    cand_src.resize(std::max<index_t>(total, 1));
    if (n_dst > 0 && total > 0) {
        compact_hits_kernel<<<grid_for(static_cast<std::size_t>(n_dst) * kCap, 256), 256, 0, s>>>(tmp.get(), count.get(), cand_off.get(), n_dst,
                                                                                               cand_src.get());
        AXIS_LAUNCH_CHECK();
        if (h_list[0] > 0)
            query_long_kernel<true, kWave, kFront0, 0><<<static_cast<unsigned>(h_list[0]), kWave, 0, s>>>(
                nodes, dst_boxes, over.get(), nullptr, cand_off.get(), cand_src.get(), level.get(), nullptr, nullptr);
        if (h_list[1] > 0)
            query_long_kernel<true, kBlock1, kFront1, 1><<<static_cast<unsigned>(h_list[1]), kBlock1, 0, s>>>(
                nodes, dst_boxes, rows1.get(), nullptr, cand_off.get(), cand_src.get(), level.get(), nullptr, nullptr);
        if (h_list[2] > 0)
            query_kernel<1><<<grid_for(h_list[2], kBlock), kBlock, 0, s>>>(nodes, src_boxes, n, dst_boxes, h_list[2], rows2.get(), nullptr,
                                                                          cand_off.get(), cand_src.get(), nullptr, nullptr, err.get());
        AXIS_LAUNCH_CHECK();
    }
    AXIS_HIP_CHECK(hipStreamSynchronize(s));
    tr.mark("query_fill");
}

void lbvh_candidates(const float *src_boxes, index_t n_src, const float *dst_boxes, index_t n_dst, DeviceBuffer<index_t> &cand_off,
                     DeviceBuffer<std::int32_t> &cand_src, hipStream_t s) {
    LbvhTree t;
    lbvh_build(src_boxes, n_src, t, s);
    lbvh_query(t, dst_boxes, n_dst, cand_off, cand_src, s);
}

}  // namespace axis_hip::detail
