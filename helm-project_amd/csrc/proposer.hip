// proposer.hip -- learned candidate proposal for the conservative broad phase.
//
// The network (mlp.hpp) proposes, per destination cell j, a search cap
// (centre c_j, radius R_j).  The candidates of j are then the source cells
// whose centroid lies in that cap and whose float box intersects j's (the
// LBVH's own test), found through a uniform 3-D bucket grid over source
// centroids.  A proposal is used only when it provably contains every source
// the plain method could clip against j:
//
//   a source i passes the reference cap test against j only if
//   angle(c_dst_j, s_i) <= r_dst_j + r_i (plus the tie band the overlap
//   kernel sends to the host), and then its centroid lies within
//   r_dst_j + r_max of c_dst_j, so r_i <= rho_j = max bucket radius over the
//   buckets meeting that ball.  Hence angle(c_j, s_i) <= d_j + r_dst_j + rho_j
//   with d_j = angle(c_j, c_dst_j), and R_j >= d_j + r_dst_j + rho_j + 1e-9
//   guarantees s_i is in the cap.
//
// Cells whose proposal fails this guard, whose cap would need more than
// kMaxSpan buckets per axis, or whose clipped overlaps do not sum to the cell
// area within 1e-12 (checked by the caller) are redone with the LBVH, so the
// weights equal the plain method's.  The same guard radius at c_j = c_dst_j
// is the training label.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>

#include "axis_hip.hpp"
#include "host_util.hpp"
#include "mlp.hpp"
#include "proposer.hpp"

namespace axis_hip::detail {

namespace {

constexpr double kLo = -1.05;           // bucket grid covers [-1.05, 1.05]^3
constexpr double kSpanLen = 2.1;
constexpr int kMaxSpan = 16;            // buckets per axis a single cap may touch
constexpr double kGuardMargin = 1e-9;   // radians, covers the cap test's tie band and rounding

__device__ inline bool excluded(const std::int32_t *flags, index_t c) { return (flags[c] & 7) != 0; }

__global__ void radius_stats_kernel(const double *cap, const std::int32_t *flags, index_t n, double *sum, unsigned long long *max_bits) {
    const index_t i = static_cast<index_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    double r = 0.0;
    if (i < n && !excluded(flags, i)) r = cap[4 * i + 3];
    double s = r, m = r;
    for (int d = 32; d >= 1; d >>= 1) {
        s += __shfl_xor(s, d, 64);
        m = fmax(m, __shfl_xor(m, d, 64));
    }
    if ((threadIdx.x & 63) == 0) {
        atomicAdd(sum, s);
        atomicMax(max_bits, static_cast<unsigned long long>(__double_as_longlong(m)));  // non-negative doubles order as integers
    }
}

struct Grid {
    int G = 1;
    double h = kSpanLen;
    const std::int32_t *beg = nullptr, *end = nullptr;  // [G^3] source ranges in `ids`
    const float *rad = nullptr;                         // [G^3] max source radius
    const std::uint32_t *ids = nullptr;
    double r_max = 0.0;
};

__device__ inline int cell_of(double v, const Grid &g) {
    int k = static_cast<int>(floor((v - kLo) / g.h));
    return k < 0 ? 0 : (k >= g.G ? g.G - 1 : k);
}

__global__ void bucket_key_kernel(const double *cap, const std::int32_t *flags, index_t n, Grid g, std::uint32_t *key, std::uint32_t *id) {
    const index_t i = static_cast<index_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const std::uint32_t G = static_cast<std::uint32_t>(g.G);
    std::uint32_t k = G * G * G;  // excluded cells sort last and belong to no bucket
    if (!excluded(flags, i)) {
        const std::uint32_t x = cell_of(cap[4 * i], g), y = cell_of(cap[4 * i + 1], g), z = cell_of(cap[4 * i + 2], g);
        k = (z * G + y) * G + x;
    }
    key[i] = k;
    id[i] = static_cast<std::uint32_t>(i);
}

__global__ void bucket_range_kernel(const std::uint32_t *key, const std::uint32_t *id, index_t n, std::uint32_t n_buckets, const double *cap,
                                    std::int32_t *beg, std::int32_t *end, float *rad) {
    const index_t p = static_cast<index_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const std::uint32_t k = key[p];
    if (k >= n_buckets) return;
    if (p == 0 || key[p - 1] != k) {
        index_t q = p;
        double r = 0.0;
        while (q < n && key[q] == k) r = fmax(r, cap[4 * static_cast<index_t>(id[q]) + 3]), ++q;
        beg[k] = static_cast<std::int32_t>(p);
        end[k] = static_cast<std::int32_t>(q);
        rad[k] = static_cast<float>(r * (1.0 + 1e-6)) + 1e-12f;  // rounded up
    }
}

__device__ inline double angle_between(const double *a, const double *b) {
    const double cx = a[1] * b[2] - a[2] * b[1], cy = a[2] * b[0] - a[0] * b[2], cz = a[0] * b[1] - a[1] * b[0];
    return atan2(sqrt(cx * cx + cy * cy + cz * cz), a[0] * b[0] + a[1] * b[1] + a[2] * b[2]);
}

// bucket index range of the ball of chord radius `ch` around p; false if too wide
__device__ inline bool ball_range(const double *p, double ch, const Grid &g, int lo[3], int hi[3]) {
    for (int a = 0; a < 3; ++a) {
        lo[a] = cell_of(p[a] - ch, g);
        hi[a] = cell_of(p[a] + ch, g);
        if (hi[a] - lo[a] + 1 > kMaxSpan) return false;
    }
    return true;
}

__device__ inline double chord(double ang) { return 2.0 * sin(fmin(ang, 3.14159265358979) * 0.5) * (1.0 + 1e-12) + 1e-15; }

// Guard radius at centre c_dst: r_dst + rho + margin; < 0 when the ball is too wide.
__device__ inline double guard_radius(const double *cd, double r_dst, const Grid &g) {
    int lo[3], hi[3];
    if (!ball_range(cd, chord(r_dst + g.r_max), g, lo, hi)) return -1.0;
    float rho = 0.0f;
    for (int z = lo[2]; z <= hi[2]; ++z)
        for (int y = lo[1]; y <= hi[1]; ++y)
            for (int x = lo[0]; x <= hi[0]; ++x) rho = fmaxf(rho, g.rad[(static_cast<index_t>(z) * g.G + y) * g.G + x]);
    return r_dst + static_cast<double>(rho) + kGuardMargin;
}

__device__ inline bool box_overlap(const float *a, const float *b) {
    return !(a[3] < b[0] || b[3] < a[0] || a[4] < b[1] || b[4] < a[1] || a[5] < b[2] || b[5] < a[2]);
}

__global__ void feature_kernel(const double *cap, const float *box, index_t n, float inv_scale, float *x) {
    const index_t j = static_cast<index_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (j >= n) return;
    for (int a = 0; a < 3; ++a) {
        x[6 * j + a] = static_cast<float>(cap[4 * j + a]);
        x[6 * j + 3 + a] = (box[6 * j + 3 + a] - box[6 * j + a]) * inv_scale;
    }
}

__global__ void label_kernel(const double *dcap, const std::int32_t *dflags, index_t n, Grid g, float scale, float *t, int *n_unlabelled) {
    const index_t j = static_cast<index_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (j >= n) return;
    const double *cd = dcap + 4 * j;
    double R = excluded(dflags, j) ? -1.0 : guard_radius(cd, cd[3], g);
    if (R <= 0.0) {
        atomicAdd(n_unlabelled, 1);
        R = cd[3] > 0.0 ? cd[3] : static_cast<double>(scale);
    }
    t[4 * j] = t[4 * j + 1] = t[4 * j + 2] = 0.0f;
    t[4 * j + 3] = static_cast<float>(log(R / scale));
}

// Per destination cell: decode the proposal, apply the guard, count the
// candidates (kFill == false) or write them (kFill == true).
template <bool kFill>
__global__ void propose_kernel(const double *dcap, const std::int32_t *dflags, const float *dbox, const float *out, index_t n, float scale,
                               const double *scap, const float *sbox, Grid g, std::int32_t *accepted, index_t *count, const index_t *off,
                               std::int32_t *cand) {
    const index_t j = static_cast<index_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (j >= n) return;
    index_t cnt = 0;
    if (!kFill) accepted[j] = 0;
    if (excluded(dflags, j)) {  // such cells never receive entries
        if (!kFill) {
            accepted[j] = 1;
            count[j] = 0;
        }
        return;
    }
    if (kFill && !accepted[j]) return;
    const double *cd = dcap + 4 * j;
    const float *o = out + 4 * j;
    double c[3] = {cd[0] + scale * static_cast<double>(o[0]), cd[1] + scale * static_cast<double>(o[1]), cd[2] + scale * static_cast<double>(o[2])};
    const double len = sqrt(c[0] * c[0] + c[1] * c[1] + c[2] * c[2]);
    if (!(len > 0.0)) {
        if (!kFill) count[j] = 0;
        return;
    }
    for (int a = 0; a < 3; ++a) c[a] /= len;
    const double R = static_cast<double>(scale) * exp(static_cast<double>(o[3]));
    if (!kFill) {
        const double gr = guard_radius(cd, cd[3], g);
        const double need = gr < 0.0 ? INFINITY : angle_between(c, cd) + (gr - cd[3]) + cd[3];
        if (!(R >= need) || !(R < 1.0)) {
            count[j] = 0;
            return;
        }
    }
    int lo[3], hi[3];
    if (!ball_range(c, chord(R), g, lo, hi)) {
        if (!kFill) count[j] = 0;
        return;
    }
    const double cosR = cos(R);
    const float *bd = dbox + 6 * j;
    index_t w = kFill ? off[j] : 0;
    for (int z = lo[2]; z <= hi[2]; ++z)
        for (int y = lo[1]; y <= hi[1]; ++y)
            for (int x = lo[0]; x <= hi[0]; ++x) {
                const index_t b = (static_cast<index_t>(z) * g.G + y) * g.G + x;
                for (int q = g.beg[b]; q < g.end[b]; ++q) {
                    const index_t i = g.ids[q];
                    const double *si = scap + 4 * i;
                    if (c[0] * si[0] + c[1] * si[1] + c[2] * si[2] < cosR) continue;
                    if (!box_overlap(sbox + 6 * i, bd)) continue;
                    if (kFill)
                        cand[w++] = static_cast<std::int32_t>(i);
                    else
                        ++cnt;
                }
            }
    if (!kFill) {
        accepted[j] = 1;
        count[j] = cnt;
    }
}

struct BucketGrid {
    Grid g;
    DeviceBuffer<std::int32_t> beg, end;
    DeviceBuffer<float> rad;
    DeviceBuffer<std::uint32_t> ids;
};

void build_grid(const MeshGeom &sg, BucketGrid &B, hipStream_t s) {
    const index_t n = sg.n_cells;
    DeviceBuffer<double> sum(1);
    DeviceBuffer<unsigned long long> mx(1);
    AXIS_HIP_CHECK(hipMemsetAsync(sum.get(), 0, sizeof(double), s));
    AXIS_HIP_CHECK(hipMemsetAsync(mx.get(), 0, sizeof(unsigned long long), s));
    if (n > 0) {
        radius_stats_kernel<<<grid_for(n, 256), 256, 0, s>>>(sg.d_cap.get(), sg.d_flags.get(), n, sum.get(), mx.get());
        AXIS_LAUNCH_CHECK();
    }
    double h_sum = 0.0;
    unsigned long long h_mx = 0;
    AXIS_HIP_CHECK(hipMemcpyAsync(&h_sum, sum.get(), sizeof(double), hipMemcpyDeviceToHost, s));
    AXIS_HIP_CHECK(hipMemcpyAsync(&h_mx, mx.get(), sizeof(unsigned long long), hipMemcpyDeviceToHost, s));
    AXIS_HIP_CHECK(hipStreamSynchronize(s));
    double r_max;
    std::memcpy(&r_max, &h_mx, sizeof r_max);
    const double r_mean = n > 0 ? h_sum / static_cast<double>(n) : 1.0;
    // buckets about one mean cell diameter wide, at most 256 per axis
    int G = static_cast<int>(std::ceil(kSpanLen / std::max(2.0 * r_mean, 1e-6)));
    G = std::clamp(G, 4, 256);
    B.g.G = G;
    B.g.h = kSpanLen / G;
    B.g.r_max = r_max * (1.0 + 1e-9);
    const std::uint32_t nb = static_cast<std::uint32_t>(G) * G * G;
    B.beg.resize(nb);
    B.end.resize(nb);
    B.rad.resize(nb);
    AXIS_HIP_CHECK(hipMemsetAsync(B.beg.get(), 0, sizeof(std::int32_t) * nb, s));
    AXIS_HIP_CHECK(hipMemsetAsync(B.end.get(), 0, sizeof(std::int32_t) * nb, s));
    AXIS_HIP_CHECK(hipMemsetAsync(B.rad.get(), 0, sizeof(float) * nb, s));
    B.ids.resize(std::max<index_t>(n, 1));
    DeviceBuffer<std::uint32_t> key(std::max<index_t>(n, 1));
    if (n > 0) {
        bucket_key_kernel<<<grid_for(n, 256), 256, 0, s>>>(sg.d_cap.get(), sg.d_flags.get(), n, B.g, key.get(), B.ids.get());
        AXIS_LAUNCH_CHECK();
        int bits = 1;
        while ((std::uint64_t(1) << bits) <= nb) ++bits;
        radix_sort_pairs_u32(key.get(), B.ids.get(), static_cast<std::size_t>(n), bits, s);
        bucket_range_kernel<<<grid_for(n, 256), 256, 0, s>>>(key.get(), B.ids.get(), n, nb, sg.d_cap.get(), B.beg.get(), B.end.get(), B.rad.get());
        AXIS_LAUNCH_CHECK();
    }
    AXIS_HIP_CHECK(hipStreamSynchronize(s));
    B.g.beg = B.beg.get();
    B.g.end = B.end.get();
    B.g.rad = B.rad.get();
    B.g.ids = B.ids.get();
}

}  // namespace

void proposer_candidates(const MeshGeom &sg, const MeshGeom &dg, const mlp::Model &model, DeviceBuffer<index_t> &cand_off,
                         DeviceBuffer<std::int32_t> &cand_src, DeviceBuffer<std::int32_t> &accepted, hipStream_t s) {
    const index_t n = dg.n_cells;
    PhaseTrace tr("propose", s);
    BucketGrid B;
    build_grid(sg, B, s);
    tr.mark("bucket_grid");
    const float scale = model.feature_scale();
    DeviceBuffer<float> x(6 * std::max<index_t>(n, 1)), out(4 * std::max<index_t>(n, 1));
    accepted.resize(std::max<index_t>(n, 1));
    cand_off.resize(n + 1);
    DeviceBuffer<index_t> count(std::max<index_t>(n, 1));
    if (n > 0) {
        feature_kernel<<<grid_for(n, 256), 256, 0, s>>>(dg.d_cap.get(), dg.d_box.get(), n, 1.0f / scale, x.get());
        AXIS_LAUNCH_CHECK();
        model.forward(x.get(), n, out.get(), s);
        tr.mark("features_mlp");
        propose_kernel<false><<<grid_for(n, 256), 256, 0, s>>>(dg.d_cap.get(), dg.d_flags.get(), dg.d_box.get(), out.get(), n, scale,
                                                              sg.d_cap.get(), sg.d_box.get(), B.g, accepted.get(), count.get(), nullptr, nullptr);
        AXIS_LAUNCH_CHECK();
    }
    exclusive_scan_i64(count.get(), cand_off.get(), n, s);
    index_t total = 0;
    AXIS_HIP_CHECK(hipMemcpyAsync(&total, cand_off.get() + n, sizeof(index_t), hipMemcpyDeviceToHost, s));
    AXIS_HIP_CHECK(hipStreamSynchronize(s));
    tr.mark("guard_count");
    if (total >= (index_t(1) << 31)) throw RuntimeError("proposer: candidate pair count exceeds 2^31");
    cand_src.resize(std::max<index_t>(total, 1));
    if (n > 0 && total > 0) {
        propose_kernel<true><<<grid_for(n, 256), 256, 0, s>>>(dg.d_cap.get(), dg.d_flags.get(), dg.d_box.get(), out.get(), n, scale, sg.d_cap.get(),
                                                             sg.d_box.get(), B.g, accepted.get(), nullptr, cand_off.get(), cand_src.get());
        AXIS_LAUNCH_CHECK();
    }
    AXIS_HIP_CHECK(hipStreamSynchronize(s));
    tr.mark("fill");
}

void proposer_training_data(const Mesh &src, const Mesh &dst, float feature_scale, float *x, float *t, hipStream_t s, index_t *n_unlabelled) {
    if (src.coord_system() == CoordinateSystem::Cartesian3D || dst.coord_system() == CoordinateSystem::Cartesian3D)
        throw InvalidArgument("proposer_training_data: spherical meshes only");
    MeshGeom sg, dg;
    GridAxes sax = unique_axes(src, s), dax = unique_axes(dst, s);
    regular_from_axes(src, sax, sg);
    regular_from_axes(dst, dax, dg);
    prepare_geometry(src, sg, s);
    prepare_geometry(dst, dg, s);
    BucketGrid B;
    build_grid(sg, B, s);
    const index_t n = dg.n_cells;
    DeviceBuffer<int> bad(1);
    AXIS_HIP_CHECK(hipMemsetAsync(bad.get(), 0, sizeof(int), s));
    if (n > 0) {
        feature_kernel<<<grid_for(n, 256), 256, 0, s>>>(dg.d_cap.get(), dg.d_box.get(), n, 1.0f / feature_scale, x);
        AXIS_LAUNCH_CHECK();
        label_kernel<<<grid_for(n, 256), 256, 0, s>>>(dg.d_cap.get(), dg.d_flags.get(), n, B.g, feature_scale, t, bad.get());
        AXIS_LAUNCH_CHECK();
    }
    int h_bad = 0;
    AXIS_HIP_CHECK(hipMemcpyAsync(&h_bad, bad.get(), sizeof(int), hipMemcpyDeviceToHost, s));
    AXIS_HIP_CHECK(hipStreamSynchronize(s));
    if (n_unlabelled) *n_unlabelled = h_bad;
}

}  // namespace axis_hip::detail
