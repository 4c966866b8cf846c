// c_abi.cpp -- extern "C" boundary of libaxis_hip.so (declared in include/axis_hip.h).
//
// Conventions of the reference interop layer (libs/axis/src/fortran/
// axis_c_interop.cpp:25-62, handle_registry.hpp:16-100): every entry point
// catches everything and returns AXIS_SUCCESS (0) / AXIS_ERROR (-1); handles are
// int tokens from a mutex-guarded registry, token 0 is invalid and tokens are
// never reused.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <unordered_map>
#include <variant>
#include <vector>

#include "../../include/axis_hip.h"
#include "axis_hip.hpp"
#include "grid_descriptor.hpp"
#include "mlp.hpp"
#include "proposer.hpp"

namespace {

using namespace axis_hip;

thread_local std::string g_last_error;

struct Registry {
    std::mutex mu;
    int next = 1;
    std::unordered_map<int, std::variant<std::shared_ptr<Mesh>, std::shared_ptr<InterpolationMatrix>, std::shared_ptr<BatchPlan>,
                                         std::shared_ptr<mlp::Model>, std::shared_ptr<HaloPattern>>>
        items;
    static Registry &get() {
        static Registry r;
        return r;
    }
    template <class T>
    int add(std::shared_ptr<T> p) {
        std::lock_guard<std::mutex> g(mu);
        int h = next++;
        items.emplace(h, std::move(p));
        return h;
    }
    template <class T>
    std::shared_ptr<T> find(int h) {
        std::lock_guard<std::mutex> g(mu);
        auto it = items.find(h);
        if (it == items.end()) throw InvalidArgument("invalid handle " + std::to_string(h));
        auto *p = std::get_if<std::shared_ptr<T>>(&it->second);
        if (!p) throw InvalidArgument("handle " + std::to_string(h) + " has the wrong kind");
        return *p;
    }
    template <class T>
    void release(int h) {
        std::lock_guard<std::mutex> g(mu);
        auto it = items.find(h);
        if (it == items.end() || !std::holds_alternative<std::shared_ptr<T>>(it->second))
            throw InvalidArgument("invalid handle " + std::to_string(h));
        items.erase(it);
    }
};

template <class F>
int guarded(F &&f) {
    try {
        f();
        return AXIS_SUCCESS;
    } catch (const std::exception &e) {
        g_last_error = e.what();
    } catch (...) {
        g_last_error = "unknown error";
    }
    return AXIS_ERROR;
}

bool is_device_ptr(const void *p) {
    if (!p) return false;
    hipPointerAttribute_t attr{};
    hipError_t e = hipPointerGetAttributes(&attr, p);
    if (e != hipSuccess) {
        (void)hipGetLastError();  // plain pageable host memory
        return false;
    }
    return attr.type == hipMemoryTypeDevice;
}

// Stages a host array into device memory when needed.
template <class T>
struct DevView {
    const T *ptr = nullptr;
    DeviceBuffer<T> own;
    DevView(const T *p, index_t n, hipStream_t s) {
        if (is_device_ptr(p) || n == 0) {
            ptr = p;
        } else {
            own.resize(n);
            own.upload(p, n, s);
            ptr = own.get();
        }
    }
};

FieldType ftype(int code) {
    switch (code) {
        case AXIS_F64: return FieldType::F64;
        case AXIS_F32: return FieldType::F32;
        case AXIS_BF16: return FieldType::BF16;
        default: throw InvalidArgument("unknown field dtype code " + std::to_string(code));
    }
}

ApplyMode amode(int code) {
    if (code == AXIS_APPLY_PLAIN) return ApplyMode::Plain;
    if (code == AXIS_APPLY_MIXED) return ApplyMode::Mixed;
    if (code == AXIS_APPLY_MIXED_ACC64) return ApplyMode::MixedAcc64;
    throw InvalidArgument("unknown apply mode " + std::to_string(code));
}

template <class T>
void copy_out(T *dst, const DeviceBuffer<T> &src, index_t n) {
    if (!dst || n <= 0) return;
    AXIS_HIP_CHECK(hipMemcpy(dst, src.get(), sizeof(T) * n, hipMemcpyDefault));
}

void check_device() {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0) {
        (void)hipGetLastError();
        throw RuntimeError("no HIP device available: libaxis_hip runs on the GPU only");
    }
}

RegridConfig config_of(const axis_regrid_config *config) {
    RegridConfig cfg;
    if (config->method < 0 || config->method > 5 || config->norm_type < 0 || config->norm_type > 1 || config->line_type < 0 ||
        config->line_type > 1 || config->unmapped < 0 || config->unmapped > 1 || config->extrap_method < 0 || config->extrap_method > 1)
        throw InvalidArgument("RegridConfig field out of range");
    cfg.method = static_cast<InterpolationMethod>(config->method);
    cfg.norm_type = static_cast<NormType>(config->norm_type);
    cfg.line_type = static_cast<LineType>(config->line_type);
    cfg.unmapped = static_cast<UnmappedAction>(config->unmapped);
    cfg.use_limiter = config->use_limiter != 0;
    cfg.extrap_method = static_cast<ExtrapolationAction>(config->extrap_method);
    return cfg;
}

}  // namespace

extern "C" {

int axis_init_c(void) {
    return guarded([] { check_device(); });
}

const char *axis_last_error_c(void) { return g_last_error.c_str(); }

int axis_mesh_create_c(const double *node_coords, int64_t n_nodes, int64_t ndim, const int64_t *conn_offsets, int64_t n_cells,
                       const int64_t *conn_indices, int coord_sys, const double *cell_areas, const int32_t *cell_mask, int *mesh_handle) {
    return guarded([&] {
        if (!mesh_handle) throw InvalidArgument("null mesh_handle");
        if (coord_sys < 0 || coord_sys > 2) throw InvalidArgument("unknown coordinate system");
        check_device();
        auto m = std::make_shared<Mesh>(node_coords, n_nodes, ndim, conn_offsets, n_cells, conn_indices, static_cast<CoordinateSystem>(coord_sys),
                                        cell_areas, cell_mask);
        *mesh_handle = Registry::get().add(std::move(m));
    });
}

int axis_mesh_regular_latlon_c(double min_x, double max_x, double min_y, double max_y, double r_x, double r_y, int *mesh_handle) {
    return guarded([&] {
        if (!mesh_handle) throw InvalidArgument("null mesh_handle");
        check_device();
        *mesh_handle = Registry::get().add(ingest::mesh_regular_latlon(min_x, max_x, min_y, max_y, r_x, r_y));
    });
}

int axis_mesh_from_descriptor_c(const void *descriptor_ptr, int *mesh_handle) {
    return guarded([&] {
        if (!descriptor_ptr || !mesh_handle) throw InvalidArgument("null argument");
        check_device();
        *mesh_handle = Registry::get().add(ingest::mesh_from_descriptor(*static_cast<const ingest::GridDescriptor *>(descriptor_ptr)));
    });
}

int axis_mesh_from_named_c(const char *name, int *mesh_handle) {
    return guarded([&] {
        if (!name || !mesh_handle) throw InvalidArgument("null argument");
        throw InvalidArgument(std::string("MeshFactory::from_named: named grids are not built ('") + name + "')");
    });
}

int axis_mesh_arrays_c(int mesh_handle, double *node_coords, int64_t *conn_offsets, int64_t *conn_indices) {
    return guarded([&] {
        auto m = Registry::get().find<Mesh>(mesh_handle);
        if (node_coords) std::memcpy(node_coords, m->h_coords.data(), sizeof(double) * m->h_coords.size());
        if (conn_offsets) std::memcpy(conn_offsets, m->h_offsets.data(), sizeof(int64_t) * m->h_offsets.size());
        if (conn_indices && !m->h_indices.empty()) std::memcpy(conn_indices, m->h_indices.data(), sizeof(int64_t) * m->h_indices.size());
    });
}

int axis_mesh_info_c(int mesh_handle, int64_t *n_nodes, int64_t *n_cells) {
    return guarded([&] {
        auto m = Registry::get().find<Mesh>(mesh_handle);
        if (n_nodes) *n_nodes = m->n_nodes();
        if (n_cells) *n_cells = m->n_cells();
    });
}

int axis_mesh_grid_info_c(int mesh_handle, int *kind, int64_t *ninj, double *out6) {
    return guarded([&] {
        auto m = Registry::get().find<Mesh>(mesh_handle);
        check_device();
        detail::GridAxes ax = detail::unique_axes(*m, nullptr);
        detail::MeshGeom g;
        const bool rect = detail::rectilinear_from_axes(*m, ax), reg = detail::regular_from_axes(*m, ax, g);
        if (kind) *kind = reg ? 2 : (rect ? 1 : 0);
        if (ninj) {
            ninj[0] = rect ? static_cast<int64_t>(ax.ulon.size()) - 1 : 0;
            ninj[1] = rect ? static_cast<int64_t>(ax.ulat.size()) - 1 : 0;
        }
        if (out6) {
            const double v[6] = {reg ? g.lon_min : 0.0, reg ? ax.ulon.back() : 0.0, reg ? g.delta_lon : 0.0,
                                 reg ? g.lat_min : 0.0, reg ? ax.ulat.back() : 0.0, reg ? g.delta_lat : 0.0};
            for (int i = 0; i < 6; ++i) out6[i] = v[i];
        }
    });
}

int axis_mesh_cell_centroids_c(int mesh_handle, float *lon, float *lat, int64_t *n_host) {
    return guarded([&] {
        if (!lon || !lat) throw InvalidArgument("null argument");
        auto m = Registry::get().find<Mesh>(mesh_handle);
        check_device();
        const index_t n = m->n_cells();
        DeviceBuffer<float2> d(std::max<index_t>(n, 1));
        const index_t fixed = detail::cell_centroids_f32(*m, d.get(), nullptr);
        std::vector<float2> h(n);
        d.download(h.data(), n, nullptr);
        AXIS_HIP_CHECK(hipStreamSynchronize(nullptr));
        for (index_t i = 0; i < n; ++i) {
            lon[i] = h[i].x;
            lat[i] = h[i].y;
        }
        if (n_host) *n_host = fixed;
    });
}

int axis_generate_weights_config_c(int src_handle, int dst_handle, const axis_regrid_config *config, int *matrix_handle) {
    return guarded([&] {
        if (!config || !matrix_handle) throw InvalidArgument("null argument");
        auto src = Registry::get().find<Mesh>(src_handle);
        auto dst = Registry::get().find<Mesh>(dst_handle);
        const RegridConfig cfg = config_of(config);
        std::shared_ptr<InterpolationMatrix> M = WeightGenerator::generate(*src, *dst, cfg, nullptr);
        *matrix_handle = Registry::get().add(std::move(M));
    });
}

int axis_generate_weights_distributed_c(int src_handle, int dst_handle, const axis_regrid_config *config, const int64_t *src_global_ids,
                                        int64_t n_src_ids, const int64_t *dst_global_ids, int64_t n_dst_ids, const int32_t *owner_of_src,
                                        int64_t n_owner, int *matrix_handle, int *pattern_handle) {
    return guarded([&] {
        if (!matrix_handle || !pattern_handle || (n_src_ids > 0 && !src_global_ids) || (n_owner > 0 && !owner_of_src))
            throw InvalidArgument("null argument");
        (void)dst_global_ids;  // carried by the reference's signature, not read by it
        (void)n_dst_ids;
        auto src = Registry::get().find<Mesh>(src_handle);
        auto dst = Registry::get().find<Mesh>(dst_handle);
        const RegridConfig cfg = config ? config_of(config) : RegridConfig{};
        auto pattern = std::make_shared<HaloPattern>();
        std::shared_ptr<InterpolationMatrix> M =
            detail::generate_distributed(*src, *dst, cfg, src_global_ids, n_src_ids, owner_of_src, n_owner, *pattern, nullptr);
        *matrix_handle = Registry::get().add(std::move(M));
        *pattern_handle = Registry::get().add(std::move(pattern));
    });
}

int axis_halo_pattern_info_c(int pattern_handle, int64_t *num_remote, int64_t *num_ranks) {
    return guarded([&] {
        auto p = Registry::get().find<HaloPattern>(pattern_handle);
        if (num_remote) *num_remote = p->num_remote();
        if (num_ranks) *num_ranks = p->num_ranks();
    });
}

int axis_halo_pattern_arrays_c(int pattern_handle, int32_t *source_ranks, int64_t *rank_offsets, int64_t *needed_global_src_ids,
                               int64_t *gather_slot) {
    return guarded([&] {
        auto p = Registry::get().find<HaloPattern>(pattern_handle);
        if (source_ranks) std::copy(p->source_ranks.begin(), p->source_ranks.end(), source_ranks);
        if (rank_offsets) std::copy(p->rank_offsets.begin(), p->rank_offsets.end(), rank_offsets);
        if (needed_global_src_ids) std::copy(p->needed_global_src_ids.begin(), p->needed_global_src_ids.end(), needed_global_src_ids);
        if (gather_slot) std::copy(p->gather_slot.begin(), p->gather_slot.end(), gather_slot);
    });
}

int axis_destroy_halo_pattern_c(int pattern_handle) {
    return guarded([&] { Registry::get().release<HaloPattern>(pattern_handle); });
}

int axis_apply_distributed_c(int matrix_handle, int pattern_handle, const double *local_src, int64_t n_local, const double *gathered_halo_src,
                             int64_t n_halo, double *dst, int64_t n_dst, int64_t n_fields, void *stream) {
    return guarded([&] {
        auto M = Registry::get().find<InterpolationMatrix>(matrix_handle);
        auto p = Registry::get().find<HaloPattern>(pattern_handle);
        check_device();
        if ((!local_src && n_local > 0) || (!gathered_halo_src && n_halo > 0) || (!dst && n_dst > 0)) throw InvalidArgument("null argument");
        detail::apply_distributed(*M, *p, local_src, n_local, gathered_halo_src, n_halo, dst, n_dst, n_fields, static_cast<hipStream_t>(stream));
    });
}

int axis_apply_distributed_callback_c(int matrix_handle, int pattern_handle, const double *local_src, int64_t n_local, axis_gather_fn gather,
                                      void *user, double *dst, int64_t n_dst, void *stream) {
    return guarded([&] {
        auto M = Registry::get().find<InterpolationMatrix>(matrix_handle);
        auto p = Registry::get().find<HaloPattern>(pattern_handle);
        check_device();
        if (!gather) throw InvalidArgument("null gather callback");
        if ((!local_src && n_local > 0) || (!dst && n_dst > 0)) throw InvalidArgument("null argument");
        // the callback is invoked exactly once with the pattern (apply.hpp:233-248)
        const index_t n = p->num_remote();
        std::vector<double> halo(std::max<index_t>(n, 1));
        const int rc = gather(p->needed_global_src_ids.data(), n, halo.data(), user);
        if (rc != 0) throw InvalidArgument("axis::solver::apply (callback): gather failed with code " + std::to_string(rc));
        const hipStream_t s = static_cast<hipStream_t>(stream);
        DeviceBuffer<double> d_halo(std::max<index_t>(n, 1));
        d_halo.upload(halo.data(), n, s);
        detail::apply_distributed(*M, *p, local_src, n_local, d_halo.get(), n, dst, n_dst, 1, s);
    });
}

int axis_generate_weights_proposed_c(int src_handle, int dst_handle, const axis_regrid_config *config, int mlp_handle, double area_tol,
                                     int *matrix_handle, int64_t *stats5) {
    return guarded([&] {
        if (!matrix_handle) throw InvalidArgument("null argument");
        auto src = Registry::get().find<Mesh>(src_handle);
        auto dst = Registry::get().find<Mesh>(dst_handle);
        auto net = Registry::get().find<mlp::Model>(mlp_handle);
        RegridConfig cfg;
        cfg.method = InterpolationMethod::Conservative1stOrder;
        if (config) cfg = config_of(config);
        ProposerStats st;
        std::shared_ptr<InterpolationMatrix> M = WeightGenerator::generate_proposed(*src, *dst, cfg, net.get(), &st, nullptr, area_tol);
        if (stats5) {
            stats5[0] = st.rows;
            stats5[1] = st.guard_rejected;
            stats5[2] = st.area_failed;
            stats5[3] = st.pairs_proposed;
            stats5[4] = st.pairs_fallback;
        }
        *matrix_handle = Registry::get().add(std::move(M));
    });
}

int axis_proposer_training_data_c(int src_handle, int dst_handle, float feature_scale, float *x, float *t, int64_t *n_unlabelled) {
    return guarded([&] {
        auto src = Registry::get().find<Mesh>(src_handle);
        auto dst = Registry::get().find<Mesh>(dst_handle);
        if (!(feature_scale > 0.0f)) throw InvalidArgument("feature_scale must be positive");
        if (dst->n_cells() > 0 && (!is_device_ptr(x) || !is_device_ptr(t)))
            throw InvalidArgument("axis_proposer_training_data_c: x and t must be device memory");
        index_t bad = 0;
        detail::proposer_training_data(*src, *dst, feature_scale, x, t, nullptr, &bad);
        if (n_unlabelled) *n_unlabelled = bad;
    });
}

int axis_generate_weights_c(int src_handle, int dst_handle, int method, int *matrix_handle) {
    axis_regrid_config cfg{0, 0, 1, 1, 0, 1};
    switch (method) {  // axis_c_interop.cpp:126-140
        case AXIS_METHOD_BILINEAR: cfg.method = 0; break;
        case AXIS_METHOD_NEAREST: cfg.method = 1; break;
        case AXIS_METHOD_CONSERVATIVE1: cfg.method = 4; break;
        default: g_last_error = "unknown method code " + std::to_string(method); return AXIS_ERROR;
    }
    return axis_generate_weights_config_c(src_handle, dst_handle, &cfg, matrix_handle);
}

int axis_generate_vector_weights_c(int src_handle, int dst_handle, const double *src_alpha, const double *dst_alpha, int method, int *matrix_u_handle,
                                   int *matrix_v_handle) {
    return guarded([&] {
        (void)Registry::get().find<Mesh>(src_handle);
        (void)Registry::get().find<Mesh>(dst_handle);
        if (!src_alpha || !dst_alpha || !matrix_u_handle || !matrix_v_handle) throw InvalidArgument("null argument");
        if (method < 0 || method > 2) throw InvalidArgument("unknown method code " + std::to_string(method));
        throw InvalidArgument("VectorWeightGenerator::generate: rotated vector weights are not built (scalar weights: axis_generate_weights_c)");
    });
}

int axis_matrix_from_coo_c(const int64_t *rows, const int64_t *cols, const double *values, int64_t nnz, int64_t n_src, int64_t n_dst,
                           const double *frac_a, const double *frac_b, const double *area_a, const double *area_b, int *matrix_handle) {
    return guarded([&] {
        if (!matrix_handle || (nnz > 0 && (!rows || !cols || !values))) throw InvalidArgument("null argument");
        check_device();
        std::shared_ptr<InterpolationMatrix> M = matrix_from_coo(rows, cols, values, nnz, n_src, n_dst, frac_a, frac_b, area_a, area_b);
        *matrix_handle = Registry::get().add(std::move(M));
    });
}

int axis_matrix_serialize_c(int matrix_handle, void *buf, uint64_t buf_size, uint64_t *bytes) {
    return guarded([&] {
        auto M = Registry::get().find<InterpolationMatrix>(matrix_handle);
        if (buf) check_device();
        const std::size_t n = serialize_axwc(*M, buf, static_cast<std::size_t>(buf_size));
        if (bytes) *bytes = n;
    });
}

int axis_matrix_deserialize_c(const void *buf, uint64_t buf_size, int *matrix_handle) {
    return guarded([&] {
        if (!matrix_handle) throw InvalidArgument("null argument");
        index_t n_src = 0, n_dst = 0, nnz = 0;
        check_axwc(buf, static_cast<std::size_t>(buf_size), n_src, n_dst, nnz);  // before touching the device
        check_device();
        std::shared_ptr<InterpolationMatrix> M = deserialize_axwc(buf, static_cast<std::size_t>(buf_size));
        *matrix_handle = Registry::get().add(std::move(M));
    });
}

int axis_matrix_write_esmf_c(int matrix_handle, const char *path) {
    return guarded([&] {
        if (!path) throw InvalidArgument("null path");
        auto M = Registry::get().find<InterpolationMatrix>(matrix_handle);
        check_device();
        write_esmf(path, *M);
    });
}

int axis_matrix_read_esmf_c(const char *path, int *matrix_handle) {
    return guarded([&] {
        if (!path || !matrix_handle) throw InvalidArgument("null argument");
        const EsmfTriplets t = parse_esmf(path);  // before touching the device
        check_device();
        std::shared_ptr<InterpolationMatrix> M = matrix_from_esmf(t);
        *matrix_handle = Registry::get().add(std::move(M));
    });
}

int axis_matrix_to_csr_c(int matrix_handle) {
    return guarded([&] { (void)Registry::get().find<InterpolationMatrix>(matrix_handle); });
}

int axis_matrix_info_c(int matrix_handle, int64_t *n_src, int64_t *n_dst, int64_t *nnz) {
    return guarded([&] {
        auto M = Registry::get().find<InterpolationMatrix>(matrix_handle);
        if (n_src) *n_src = M->n_src;
        if (n_dst) *n_dst = M->n_dst;
        if (nnz) *nnz = M->nnz;
    });
}

int axis_matrix_csr_c(int matrix_handle, int64_t *row_ptr, int64_t *col_idx, double *values) {
    return guarded([&] {
        auto M = Registry::get().find<InterpolationMatrix>(matrix_handle);
        copy_out(row_ptr, M->row_ptr, M->n_dst + 1);
        copy_out(col_idx, M->col, M->nnz);
        copy_out(values, M->val, M->nnz);
    });
}

int axis_matrix_fractions_c(int matrix_handle, double *frac_a, double *frac_b, double *area_a, double *area_b) {
    return guarded([&] {
        auto M = Registry::get().find<InterpolationMatrix>(matrix_handle);
        copy_out(frac_a, M->frac_a, M->n_src);
        copy_out(frac_b, M->frac_b, M->n_dst);
        copy_out(area_a, M->area_a, M->n_src);
        copy_out(area_b, M->area_b, M->n_dst);
    });
}

int axis_matrix_stats_c(int matrix_handle, int64_t *out4, double *gen_ms) {
    return guarded([&] {
        auto M = Registry::get().find<InterpolationMatrix>(matrix_handle);
        if (out4) {
            out4[0] = M->n_candidates;
            out4[1] = M->n_host_arbitrated;
            out4[2] = M->n_cap_borderline;
            out4[3] = M->max_row_nnz;
        }
        if (gen_ms) *gen_ms = M->gen_ms;
    });
}

int axis_matrix_phase_ms_c(int matrix_handle, double *out, int n) {
    return guarded([&] {
        auto M = Registry::get().find<InterpolationMatrix>(matrix_handle);
        if (n < 0 || (n > 0 && !out)) throw InvalidArgument("axis_matrix_phase_ms_c: bad output array");
        for (int k = 0; k < n; ++k) out[k] = k < 4 ? M->phase_ms[k] : -1.0;
    });
}

int axis_apply_c(int matrix_handle, const double *src, double *dst, int n_src, int n_dst) {
    return guarded([&] {
        auto M = Registry::get().find<InterpolationMatrix>(matrix_handle);
        // apply.hpp:63-71: validation before dst is touched
        if (static_cast<index_t>(n_src) != M->n_src)
            throw InvalidArgument("axis::solver::apply: src.extent(0) (" + std::to_string(n_src) + ") != matrix.n_src() (" +
                                  std::to_string(M->n_src) + ")");
        if (static_cast<index_t>(n_dst) != M->n_dst)
            throw InvalidArgument("axis::solver::apply: dst.extent(0) (" + std::to_string(n_dst) + ") != matrix.n_dst() (" +
                                  std::to_string(M->n_dst) + ")");
        hipStream_t s = nullptr;
        DevView<double> x(src, n_src, s);
        const bool dev_out = is_device_ptr(dst);
        DeviceBuffer<double> ybuf;
        double *y = dst;
        if (!dev_out && n_dst > 0) {
            ybuf.resize(n_dst);
            y = ybuf.get();
        }
        apply(*M, field_view<const double, 1>(x.ptr, static_cast<std::size_t>(n_src)), field_view<double, 1>(y, static_cast<std::size_t>(n_dst)), s);
        if (!dev_out && n_dst > 0) ybuf.download(dst, n_dst, s);
        AXIS_HIP_CHECK(hipStreamSynchronize(s));
    });
}

int axis_batch_apply_async_c(int matrix_handle, const void *src, int src_dtype, void *dst, int dst_dtype, int64_t n_src, int64_t n_dst,
                             int64_t n_fields, int mode, void *stream) {
    return guarded([&] {
        auto M = Registry::get().find<InterpolationMatrix>(matrix_handle);
        batch_apply(*M, src, ftype(src_dtype), dst, ftype(dst_dtype), n_src, n_dst, n_fields, amode(mode), static_cast<hipStream_t>(stream));
    });
}

int axis_batch_apply_noalloc_c(int matrix_handle, const void *src, int src_dtype, void *dst, int dst_dtype, int64_t n_src, int64_t n_dst,
                               int64_t n_fields, int mode, void *stream) {
    return guarded([&] {
        auto M = Registry::get().find<InterpolationMatrix>(matrix_handle);
        batch_apply_noalloc(*M, src, ftype(src_dtype), dst, ftype(dst_dtype), n_src, n_dst, n_fields, amode(mode), static_cast<hipStream_t>(stream));
    });
}

int axis_batch_apply_c(int matrix_handle, const void *src, int src_dtype, void *dst, int dst_dtype, int64_t n_src, int64_t n_dst,
                       int64_t n_fields, int mode, void *stream) {
    return guarded([&] {
        auto M = Registry::get().find<InterpolationMatrix>(matrix_handle);
        auto s = static_cast<hipStream_t>(stream);
        if (amode(mode) == ApplyMode::Plain && ftype(src_dtype) == FieldType::F64 && ftype(dst_dtype) == FieldType::F64 && n_src >= 0 && n_dst >= 0 &&
            n_fields >= 0) {  // the reference's batch_apply on field views (apply.hpp:336-343)
            batch_apply(*M, field_view<const double, 2>(static_cast<const double *>(src), static_cast<std::size_t>(n_src), static_cast<std::size_t>(n_fields)),
                        field_view<double, 2>(static_cast<double *>(dst), static_cast<std::size_t>(n_dst), static_cast<std::size_t>(n_fields)), s);
            return;
        }
        batch_apply(*M, src, ftype(src_dtype), dst, ftype(dst_dtype), n_src, n_dst, n_fields, amode(mode), s);
        AXIS_HIP_CHECK(hipStreamSynchronize(s));
    });
}

int axis_nan_batch_apply_c(int matrix_handle, const double *src, double *dst, int64_t n_src, int64_t n_dst, int64_t n_fields, double na_thres,
                           void *stream) {
    return guarded([&] {
        auto M = Registry::get().find<InterpolationMatrix>(matrix_handle);
        auto s = static_cast<hipStream_t>(stream);
        if (n_src < 0 || n_dst < 0 || n_fields < 0) throw InvalidArgument("axis::solver::nan_batch_apply: negative extent");
        nan_batch_apply(*M, field_view<const double, 2>(src, static_cast<std::size_t>(n_src), static_cast<std::size_t>(n_fields)),
                        field_view<double, 2>(dst, static_cast<std::size_t>(n_dst), static_cast<std::size_t>(n_fields)), na_thres, s);
    });
}

int axis_apply_masked_c(int matrix_handle, const double *src, double *dst, const int32_t *dst_mask, int64_t n_mask, int64_t n_src, int64_t n_dst,
                        int64_t n_fields, void *stream) {
    return guarded([&] {
        auto M = Registry::get().find<InterpolationMatrix>(matrix_handle);
        auto s = static_cast<hipStream_t>(stream);
        masked_apply(*M, src, dst, dst_mask, n_mask, n_src, n_dst, n_fields, s);
        AXIS_HIP_CHECK(hipStreamSynchronize(s));
    });
}

int axis_autotune_apply_c(int matrix_handle, const void *src, int src_dtype, void *dst, int dst_dtype, int64_t n_fields, int mode, void *stream,
                          int *fields_per_step, int *field_chunks, int *staged) {
    return guarded([&] {
        auto M = Registry::get().find<InterpolationMatrix>(matrix_handle);
        auto r = autotune_apply(*M, src, ftype(src_dtype), dst, ftype(dst_dtype), n_fields, amode(mode), static_cast<hipStream_t>(stream));
        if (fields_per_step) *fields_per_step = r.fields_per_step;
        if (field_chunks) *field_chunks = r.field_chunks;
        if (staged) *staged = r.wave ? 2 : (r.staged ? 1 : 0);
    });
}

int axis_apply_set_shape_c(int matrix_handle, int64_t n_fields, int src_dtype, int dst_dtype, int mode, int fields_per_step, int field_chunks,
                           int staged) {
    return guarded([&] {
        auto M = Registry::get().find<InterpolationMatrix>(matrix_handle);
        if (staged < 0 || staged > 2) throw InvalidArgument("staged must be 0, 1 or 2");
        ApplyShape sh;
        sh.fields_per_step = fields_per_step;
        sh.field_chunks = field_chunks;
        sh.staged = staged != 0;
        sh.wave = staged == 2;
        set_apply_shape(*M, n_fields, ftype(src_dtype), ftype(dst_dtype), amode(mode), sh);
    });
}

int axis_check_conservation_c(int matrix_handle, const double *src, const double *dst, int norm_type, double *out4) {
    return guarded([&] {
        if (!out4) throw InvalidArgument("null out4");
        if (norm_type < 0 || norm_type > 1) throw InvalidArgument("unknown norm type");
        auto M = Registry::get().find<InterpolationMatrix>(matrix_handle);
        hipStream_t s = nullptr;
        DevView<double> x(src, M->n_src, s), y(dst, M->n_dst, s);
        ConservationReport r = check_conservation(*M, x.ptr, y.ptr, static_cast<NormType>(norm_type), s);
        out4[0] = r.src_integral;
        out4[1] = r.dst_integral;
        out4[2] = r.absolute_error;
        out4[3] = r.relative_error;
    });
}

int axis_destroy_mesh_c(int handle) {
    return guarded([&] { Registry::get().release<Mesh>(handle); });
}

int axis_destroy_matrix_c(int handle) {
    return guarded([&] { Registry::get().release<InterpolationMatrix>(handle); });
}

int axis_batch_plan_create_c(int n_pairs, const int *matrix_handles, const void *const *src, void *const *dst, const int64_t *n_fields,
                             int src_dtype, int dst_dtype, int mode, int autotune, void *stream, int *plan_handle) {
    return guarded([&] {
        if (!plan_handle) throw InvalidArgument("null plan_handle");
        if (n_pairs <= 0 || !matrix_handles || !src || !dst || !n_fields) throw InvalidArgument("axis_batch_plan_create_c: empty or null pair list");
        if (autotune < 0 || autotune > 2) throw InvalidArgument("axis_batch_plan_create_c: autotune must be 0, 1 or 2");
        std::vector<std::shared_ptr<InterpolationMatrix>> mats;
        std::vector<const void *> xs;
        std::vector<void *> ys;
        std::vector<index_t> ks;
        for (int p = 0; p < n_pairs; ++p) {
            mats.push_back(Registry::get().find<InterpolationMatrix>(matrix_handles[p]));
            xs.push_back(src[p]);
            ys.push_back(dst[p]);
            ks.push_back(n_fields[p]);
            if (n_fields[p] > 0 && (!is_device_ptr(src[p]) || !is_device_ptr(dst[p])))
                throw InvalidArgument("axis_batch_plan_create_c: fields of pair " + std::to_string(p) + " are not device memory");
        }
        auto plan = std::make_shared<BatchPlan>(std::move(mats), std::move(xs), std::move(ys), std::move(ks), ftype(src_dtype), ftype(dst_dtype),
                                                amode(mode), autotune, static_cast<hipStream_t>(stream));
        *plan_handle = Registry::get().add(std::move(plan));
    });
}

int axis_batch_plan_run_c(int plan_handle, void *stream, int sync) {
    return guarded([&] { Registry::get().find<BatchPlan>(plan_handle)->run(static_cast<hipStream_t>(stream), sync != 0); });
}

int axis_batch_plan_info_c(int plan_handle, int *fields_per_step, int *fields_per_item, int64_t *n_items, int *n_workgroups) {
    return guarded([&] {
        auto P = Registry::get().find<BatchPlan>(plan_handle);
        const BatchShape sh = P->shape();
        if (fields_per_step) *fields_per_step = sh.fields_per_step;
        if (fields_per_item) *fields_per_item = sh.fields_per_item;
        if (n_items) *n_items = P->n_items();
        if (n_workgroups) *n_workgroups = P->n_workgroups();
    });
}

int axis_destroy_batch_plan_c(int handle) {
    return guarded([&] { Registry::get().release<BatchPlan>(handle); });
}

int axis_mlp_create_c(const float *params, uint64_t seed, float feature_scale, int *mlp_handle) {
    return guarded([&] {
        if (!mlp_handle) throw InvalidArgument("null mlp_handle");
        std::vector<float> p(mlp::kParams);
        if (params)
            std::memcpy(p.data(), params, sizeof(float) * mlp::kParams);
        else
            mlp::init_params(seed, p.data());
        *mlp_handle = Registry::get().add(std::make_shared<mlp::Model>(p.data(), feature_scale));
    });
}

int axis_mlp_params_c(int mlp_handle, float *params, float *feature_scale) {
    return guarded([&] {
        auto M = Registry::get().find<mlp::Model>(mlp_handle);
        if (params) M->get_params(params, nullptr);
        if (feature_scale) *feature_scale = M->feature_scale();
    });
}

int axis_mlp_set_params_c(int mlp_handle, const float *params) {
    return guarded([&] {
        if (!params) throw InvalidArgument("null params");
        Registry::get().find<mlp::Model>(mlp_handle)->set_params(params, nullptr);
    });
}

int axis_mlp_forward_c(int mlp_handle, const float *x, float *out, int64_t n, void *stream) {
    return guarded([&] {
        if (n < 0) throw InvalidArgument("negative sample count");
        if (n > 0 && (!is_device_ptr(x) || !is_device_ptr(out))) throw InvalidArgument("axis_mlp_forward_c: x and out must be device memory");
        auto M = Registry::get().find<mlp::Model>(mlp_handle);
        M->forward(x, n, out, static_cast<hipStream_t>(stream));
        AXIS_HIP_CHECK(hipStreamSynchronize(static_cast<hipStream_t>(stream)));
    });
}

int axis_mlp_gradient_c(int mlp_handle, const float *x, const float *t, int64_t n, float wc, float tau, float *grad, float *loss, void *stream) {
    return guarded([&] {
        if (!is_device_ptr(x) || !is_device_ptr(t) || !is_device_ptr(grad)) throw InvalidArgument("axis_mlp_gradient_c: device pointers expected");
        auto M = Registry::get().find<mlp::Model>(mlp_handle);
        const float l = M->gradient(x, t, n, wc, tau, grad, static_cast<hipStream_t>(stream));
        if (loss) *loss = l;
    });
}

int axis_mlp_train_c(int mlp_handle, const float *x, const float *t, int64_t n, int64_t batch, float lr, float wc, float tau, float *mean_loss,
                     void *stream) {
    return guarded([&] {
        if (!is_device_ptr(x) || !is_device_ptr(t)) throw InvalidArgument("axis_mlp_train_c: device pointers expected");
        auto M = Registry::get().find<mlp::Model>(mlp_handle);
        const float l = M->train(x, t, n, batch, lr, wc, tau, static_cast<hipStream_t>(stream));
        if (mean_loss) *mean_loss = l;
    });
}

int axis_destroy_mlp_c(int handle) {
    return guarded([&] { Registry::get().release<mlp::Model>(handle); });
}

}  // extern "C"
