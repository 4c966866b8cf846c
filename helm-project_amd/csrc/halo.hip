// Distributed mode of one rank, on its GPU:
//
//   * generate_distributed (WeightGenerator::generate with global ids,
//     weight_generator.cpp:4525-4680): the rank's local matrix, then the
//     HaloPattern of the off-rank source cells its nonzeros read -- grouped
//     by owning rank, each given a slot in the gathered halo buffer -- and
//     the matrix's columns remapped so that slot g reads column
//     n_local + g;
//   * apply_distributed (the gathered-buffer overload of apply,
//     apply.hpp:151-211): dst = S [local_src ; halo].
//
// The exchange itself (HALO's job in the reference) is outside: the rank
// hands in the gathered halo values.  Column marks, the remap and the
// per-row re-sort run on the GPU; the grouping by owner runs on the host
// over the (small) set of used off-rank columns.
#include <algorithm>
#include <numeric>

#include "axis_hip.hpp"

namespace axis_hip::detail {
namespace {

__global__ void mark_used_kernel(const index_t *__restrict__ col, index_t nnz, index_t n_local, std::uint8_t *__restrict__ used) {
    const index_t e = static_cast<index_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (e >= nnz) return;
    const index_t c = col[e];
    if (c >= 0 && c < n_local) used[c] = 1;  // a plain byte store: every writer stores the same value
}

// new column of every entry, and the (row, new column) sort key
__global__ void remap_kernel(const index_t *__restrict__ row_ptr, index_t n_rows, const index_t *__restrict__ col,
                             const index_t *__restrict__ remap, index_t *__restrict__ new_col, std::uint64_t *__restrict__ key,
                             std::uint32_t *__restrict__ perm) {
    const index_t r = static_cast<index_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (r >= n_rows) return;
    for (index_t e = row_ptr[r]; e < row_ptr[r + 1]; ++e) {
        const index_t c = remap[col[e]];
        new_col[e] = c;
        key[e] = (static_cast<std::uint64_t>(r) << 32) | static_cast<std::uint64_t>(c);
        perm[e] = static_cast<std::uint32_t>(e);
    }
}

__global__ void permute_kernel(const std::uint32_t *__restrict__ perm, index_t nnz, const index_t *__restrict__ col_in, const double *__restrict__ val_in,
                               index_t *__restrict__ col_out, double *__restrict__ val_out) {
    const index_t e = static_cast<index_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (e >= nnz) return;
    col_out[e] = col_in[perm[e]];
    val_out[e] = val_in[perm[e]];
}

// [local_src ; halo] for each field of a layout_left batch
__global__ void concat_kernel(const double *__restrict__ a, index_t na, const double *__restrict__ b, index_t nb, index_t k, double *__restrict__ out) {
    const index_t n = na + nb;
    const index_t i = static_cast<index_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n * k) return;
    const index_t f = i / n, c = i % n;
    out[i] = c < na ? a[f * na + c] : b[f * nb + (c - na)];
}

}  // namespace

std::unique_ptr<InterpolationMatrix> generate_distributed(const Mesh &src, const Mesh &dst, const RegridConfig &cfg, const index_t *src_global_ids,
                                                          index_t n_src_ids, const std::int32_t *owner_of_src, index_t n_owner, HaloPattern &pattern,
                                                          hipStream_t s) {
    const index_t n_local = src.n_cells();
    if (n_src_ids != n_local)
        throw InvalidArgument("WeightGenerator::generate (distributed): src_global_ids has " + std::to_string(n_src_ids) + " entries for " +
                              std::to_string(n_local) + " source cells");
    std::unique_ptr<InterpolationMatrix> M = WeightGenerator::generate(src, dst, cfg, s);
    // the rank that owns the first local source cell is this rank
    int local_rank = -1;
    if (n_local > 0 && src_global_ids[0] >= 0 && src_global_ids[0] < n_owner) local_rank = owner_of_src[src_global_ids[0]];
    // local columns the nonzeros read
    DeviceBuffer<std::uint8_t> used(std::max<index_t>(n_local, 1));
    AXIS_HIP_CHECK(hipMemsetAsync(used.get(), 0, used.bytes(), s));
    if (M->nnz > 0) {
        mark_used_kernel<<<grid_for(M->nnz, 256), 256, 0, s>>>(M->col.get(), M->nnz, n_local, used.get());
        AXIS_LAUNCH_CHECK();
    }
    std::vector<std::uint8_t> h_used(n_local);
    used.download(h_used.data(), n_local, s);
    AXIS_HIP_CHECK(hipStreamSynchronize(s));
    // off-rank columns, grouped by owner rank, ascending local column within a rank
    std::vector<index_t> off;
    for (index_t c = 0; c < n_local; ++c) {
        if (!h_used[c]) continue;
        const index_t gid = src_global_ids[c];
        if (gid < 0 || gid >= n_owner) continue;
        if (owner_of_src[gid] != local_rank) off.push_back(c);
    }
    std::stable_sort(off.begin(), off.end(), [&](index_t a, index_t b) { return owner_of_src[src_global_ids[a]] < owner_of_src[src_global_ids[b]]; });
    pattern = HaloPattern{};
    pattern.rank_offsets.clear();
    for (std::size_t i = 0; i < off.size(); ++i) {
        const int rank = owner_of_src[src_global_ids[off[i]]];
        if (pattern.source_ranks.empty() || pattern.source_ranks.back() != rank) {
            pattern.source_ranks.push_back(rank);
            pattern.rank_offsets.push_back(static_cast<index_t>(i));
        }
        pattern.needed_global_src_ids.push_back(src_global_ids[off[i]]);
        pattern.gather_slot.push_back(static_cast<index_t>(i));
    }
    pattern.rank_offsets.push_back(static_cast<index_t>(off.size()));
    const index_t n_remote = static_cast<index_t>(off.size());
    const index_t n_ext = n_local + n_remote;
    if (n_remote > 0 && M->nnz > 0) {
        if (M->n_dst >= (index_t(1) << 32) || n_ext >= (index_t(1) << 32)) throw InvalidArgument("generate (distributed): more than 2^32 rows or columns");
        std::vector<index_t> remap(n_local);
        std::iota(remap.begin(), remap.end(), index_t(0));
        for (index_t g = 0; g < n_remote; ++g) remap[off[g]] = n_local + g;
        DeviceBuffer<index_t> d_remap(n_local), new_col(M->nnz), sorted_col(M->nnz);
        DeviceBuffer<double> sorted_val(M->nnz);
        DeviceBuffer<std::uint64_t> key(M->nnz);
        DeviceBuffer<std::uint32_t> perm(M->nnz);
        d_remap.upload(remap.data(), n_local, s);
        remap_kernel<<<grid_for(M->n_dst, 256), 256, 0, s>>>(M->row_ptr.get(), M->n_dst, M->col.get(), d_remap.get(), new_col.get(), key.get(),
                                                             perm.get());
        AXIS_LAUNCH_CHECK();
        // (row, col) order again: a stable sort of the remapped keys
        int bits = 40;
        while (bits < 64 && (index_t(1) << (bits - 32)) < M->n_dst) bits += 8;
        radix_sort_pairs_u64(key.get(), perm.get(), static_cast<std::size_t>(M->nnz), bits, s);
        permute_kernel<<<grid_for(M->nnz, 256), 256, 0, s>>>(perm.get(), M->nnz, new_col.get(), M->val.get(), sorted_col.get(), sorted_val.get());
        AXIS_LAUNCH_CHECK();
        AXIS_HIP_CHECK(hipStreamSynchronize(s));
        M->col = std::move(sorted_col);
        M->val = std::move(sorted_val);
    }
    if (n_remote > 0) {
        // source-side arrays cover the extended source space; the halo slots
        // hold zero fractions and areas
        DeviceBuffer<double> fa(n_ext), aa(n_ext);
        AXIS_HIP_CHECK(hipMemsetAsync(fa.get(), 0, fa.bytes(), s));
        AXIS_HIP_CHECK(hipMemsetAsync(aa.get(), 0, aa.bytes(), s));
        if (n_local > 0) {
            AXIS_HIP_CHECK(hipMemcpyAsync(fa.get(), M->frac_a.get(), sizeof(double) * n_local, hipMemcpyDeviceToDevice, s));
            AXIS_HIP_CHECK(hipMemcpyAsync(aa.get(), M->area_a.get(), sizeof(double) * n_local, hipMemcpyDeviceToDevice, s));
        }
        AXIS_HIP_CHECK(hipStreamSynchronize(s));
        M->frac_a = std::move(fa);
        M->area_a = std::move(aa);
    }
    M->n_src = n_ext;
    M->reset_apply_layout();
    M->len_hist_ready = false;
    return M;
}

void apply_distributed(InterpolationMatrix &A, const HaloPattern &pattern, const double *local_src, index_t n_local, const double *halo,
                       index_t n_halo, double *dst, index_t n_dst, index_t k, hipStream_t s) {
    // extent validation before any write (apply.hpp:155-170)
    if (n_dst != A.n_dst)
        throw InvalidArgument("axis::solver::apply (distributed): dst.extent(0) (" + std::to_string(n_dst) + ") != matrix.n_dst() (" +
                              std::to_string(A.n_dst) + ")");
    const index_t n_remote = pattern.num_remote();
    if (n_halo != n_remote)
        throw InvalidArgument("axis::solver::apply (distributed): gathered_halo_src.extent(0) (" + std::to_string(n_halo) +
                              ") != pattern.num_remote() (" + std::to_string(n_remote) + ")");
    if (n_local + n_remote != A.n_src)
        throw InvalidArgument("axis::solver::apply (distributed): local_src.extent(0) (" + std::to_string(n_local) + ") + pattern.num_remote() (" +
                              std::to_string(n_remote) + ") != matrix.n_src() (" + std::to_string(A.n_src) + ")");
    if (k < 1) throw InvalidArgument("axis::solver::apply (distributed): n_fields < 1");
    const index_t n = A.n_src;
    DeviceBuffer<double> x(std::max<index_t>(n * k, 1));
    if (n * k > 0) {
        concat_kernel<<<grid_for(n * k, 256), 256, 0, s>>>(local_src, n_local, halo, n_halo, k, x.get());
        AXIS_LAUNCH_CHECK();
    }
    batch_apply(A, x.get(), FieldType::F64, dst, FieldType::F64, n, n_dst, k, ApplyMode::Plain, s);
    AXIS_HIP_CHECK(hipStreamSynchronize(s));  // x is released on return
}

}  // namespace axis_hip::detail
