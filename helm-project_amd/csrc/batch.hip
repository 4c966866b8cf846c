// batch.hip -- several regrid pairs of uneven sizes applied in one
// persistent, work-queue launch captured in a HIP graph.
//
// The reference applies one matrix per call (batch_apply,
// libs/axis/include/axis/solver/apply.hpp:275-343); a model coupling step
// applies many (levels x variables for several grid pairs).  A BatchPlan
// binds a list of (matrix, src fields, dst fields, k) and builds:
//   * work items (pair, tile, field range), the biggest first, split into
//     eight queues, one per XCD, each a contiguous run of tiles per field
//     range so that concurrently running tiles share source columns in that
//     XCD's L2;
//   * a persistent kernel sized to fill the GPU: each workgroup pops items
//     from its XCD's queue (one atomic per item), then steals from the other
//     queues, and runs the staged-tile apply of apply_kernels.hpp on each;
//   * the long-row reductions of every pair;
// all recorded once into a hipGraph (queue counters are cleared by a kernel
// inside the graph, not by a memset node) and replayed by run().  An
// autotuner picks the fields per step and the fields per work item by
// timing graph replays.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <numeric>

#include "apply_kernels.hpp"
#include "axis_hip.hpp"

namespace axis_hip {

namespace {

using apply_detail::kTileSeg;
using apply_detail::kTileCols;
using apply_detail::Mat;
using apply_detail::NumTraits;

constexpr int kQueues = 8;  // one per XCD

struct PairDesc {
    Mat A;
    const void *X;
    void *Y;
    double *P;
};

struct Item {
    std::int32_t pair, tile, f0, f1;
};

__global__ void clear_counters_kernel(unsigned *ctr) {
    if (threadIdx.x < kQueues) ctr[threadIdx.x] = 0u;
}

template <int E, int FU, int NUM, class XT, class YT>
// at least 5 waves per SIMD: every workgroup stays resident for the whole
// launch, so occupancy is fixed at launch time
__global__ void __launch_bounds__(kTileSeg, 5) persistent_apply_kernel(const PairDesc *pairs, const Item *items, const int *qbeg, unsigned *ctr) {
    __shared__ apply_detail::Stage<FU, NUM> S;
    __shared__ int s_item;
    const int home = static_cast<int>(blockIdx.x & (kQueues - 1));  // workgroups are dealt to XCDs round-robin
    for (int qq = 0; qq < kQueues; ++qq) {
        const int q = (home + qq) & (kQueues - 1);
        const int b = qbeg[q], e = qbeg[q + 1];
        if (b >= e) continue;
        while (true) {
            if (threadIdx.x == 0) {
                const unsigned got = atomicAdd(&ctr[q], 1u);
                s_item = got < static_cast<unsigned>(e - b) ? b + static_cast<int>(got) : -1;
            }
            __syncthreads();
            const int it = s_item;
            __syncthreads();
            if (it < 0) break;  // queue q drained (whole workgroup)
            const Item I = items[it];
            // the pair's operands copied into registers once per item (read
            // through the pointer, every use inside the field loop would be
            // reloaded: the kernel's stores could alias them)
            const PairDesc D = pairs[I.pair];
            apply_detail::tile_apply<E, FU, NUM, XT, YT>(D.A, static_cast<const XT *>(D.X), static_cast<YT *>(D.Y), D.P, I.tile, I.f0, I.f1,
                                                         S);
        }
    }
}

template <class YT>
__global__ void pair_long_reduce_kernel(const PairDesc *pairs, int p, index_t k) {
    const PairDesc &D = pairs[p];
    const index_t t = static_cast<index_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= static_cast<index_t>(D.A.n_long) * k) return;
    apply_detail::long_reduce(D.A, D.P, static_cast<YT *>(D.Y), k, t);
}

}  // namespace

struct BatchPlan::Impl {
    std::vector<std::shared_ptr<InterpolationMatrix>> mats;
    std::vector<const void *> src;
    std::vector<void *> dst;
    std::vector<index_t> k;
    FieldType xt, yt;
    ApplyMode mode;
    int max_e = 4;
    DeviceBuffer<PairDesc> d_pairs;
    DeviceBuffer<Item> d_items;
    DeviceBuffer<int> d_qbeg;
    DeviceBuffer<unsigned> d_ctr;
    std::vector<DeviceBuffer<double>> partial;  // long-row partial sums of each pair, owned by the plan
    hipStream_t cap = nullptr;                  // private stream the graph is recorded on
    index_t n_items = 0;
    int n_wg = 0;
    BatchShape shape;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;

    ~Impl() {
        drop_graph();
        if (cap) (void)hipStreamDestroy(cap);
    }
    void drop_graph() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        exec = nullptr;
        graph = nullptr;
    }

    // Work items for `fields_per_item`: (pair, tile, field range), the
    // largest first (longest-processing-time order), then cut into kQueues
    // contiguous runs of nearly equal estimated cost.
    void build_items(int fields_per_item, hipStream_t s) {
        struct Run {
            int pair;
            index_t f0, f1;
            double cost_per_tile;
        };
        std::vector<Run> runs;
        for (std::size_t p = 0; p < mats.size(); ++p) {
            const auto &A = *mats[p];
            const double per_tile = (A.n_tiles > 0 ? static_cast<double>(A.nnz + A.n_src) / static_cast<double>(A.n_tiles) : 0.0) + 256.0;
            for (index_t f0 = 0; f0 < k[p]; f0 += fields_per_item)
                runs.push_back({static_cast<int>(p), f0, std::min<index_t>(k[p], f0 + fields_per_item), per_tile});
        }
        std::stable_sort(runs.begin(), runs.end(), [](const Run &a, const Run &b) {
            return a.cost_per_tile * static_cast<double>(a.f1 - a.f0) > b.cost_per_tile * static_cast<double>(b.f1 - b.f0);
        });
        std::vector<Item> h;
        std::vector<double> cost;
        for (const Run &r : runs) {
            const index_t nt = mats[r.pair]->n_tiles;
            for (index_t t = 0; t < nt; ++t) {
                h.push_back(Item{r.pair, static_cast<std::int32_t>(t), static_cast<std::int32_t>(r.f0), static_cast<std::int32_t>(r.f1)});
                cost.push_back(r.cost_per_tile * static_cast<double>(r.f1 - r.f0));
            }
        }
        if (h.size() >= (std::size_t(1) << 31)) throw InvalidArgument("BatchPlan: too many work items");
        n_items = static_cast<index_t>(h.size());
        // queue q takes items [qbeg[q], qbeg[q+1]): contiguous runs of equal cost
        std::vector<int> qbeg(kQueues + 1, static_cast<int>(n_items));
        qbeg[0] = 0;
        const double total = std::accumulate(cost.begin(), cost.end(), 0.0);
        double acc = 0.0;
        int q = 1;
        for (index_t i = 0; i < n_items && q < kQueues; ++i) {
            acc += cost[i];
            while (q < kQueues && acc >= total * q / kQueues) qbeg[q++] = static_cast<int>(i + 1);
        }
        d_items.resize(std::max<std::size_t>(h.size(), 1));
        if (!h.empty()) d_items.upload(h.data(), h.size(), s);
        d_qbeg.resize(kQueues + 1);
        d_qbeg.upload(qbeg.data(), kQueues + 1, s);
        AXIS_HIP_CHECK(hipStreamSynchronize(s));
    }

    template <int E, int FU, int NUM, class XT, class YT>
    void record_t(hipStream_t s) {
        auto kern = persistent_apply_kernel<E, FU, NUM, XT, YT>;
        int dev = 0, per_cu = 0;
        AXIS_HIP_CHECK(hipGetDevice(&dev));
        hipDeviceProp_t prop{};
        AXIS_HIP_CHECK(hipGetDeviceProperties(&prop, dev));
        AXIS_HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, kTileSeg, 0));
        n_wg = std::max(1, per_cu) * prop.multiProcessorCount;
        n_wg = (n_wg + kQueues - 1) / kQueues * kQueues;
        (void)s;
        AXIS_HIP_CHECK(hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal));
        clear_counters_kernel<<<1, 64, 0, cap>>>(d_ctr.get());
        if (n_items > 0) kern<<<static_cast<unsigned>(n_wg), kTileSeg, 0, cap>>>(d_pairs.get(), d_items.get(), d_qbeg.get(), d_ctr.get());
        for (std::size_t p = 0; p < mats.size(); ++p) {
            const index_t n = static_cast<index_t>(mats[p]->n_long_rows) * k[p];
            if (n > 0) pair_long_reduce_kernel<YT><<<grid_for(n, 256), 256, 0, cap>>>(d_pairs.get(), static_cast<int>(p), k[p]);
        }
        hipError_t launch_err = hipGetLastError();
        hipGraph_t g = nullptr;
        hipError_t cap_err = hipStreamEndCapture(cap, &g);
        AXIS_HIP_CHECK(launch_err);
        AXIS_HIP_CHECK(cap_err);
        graph = g;
        AXIS_HIP_CHECK(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
    }

    template <int FU, int NUM, class XT, class YT>
    void record_fu(hipStream_t s) {
        switch (max_e) {
            case 4: return record_t<4, FU, NUM, XT, YT>(s);
            case 8: return record_t<8, FU, NUM, XT, YT>(s);
            case 12: return record_t<12, FU, NUM, XT, YT>(s);
            default: return record_t<16, FU, NUM, XT, YT>(s);
        }
    }

    template <int NUM, class XT, class YT>
    void record_xy(int fu, hipStream_t s) {
        switch (fu) {
            case 1: return record_fu<1, NUM, XT, YT>(s);
            case 2: return record_fu<2, NUM, XT, YT>(s);
            default: return record_fu<4, NUM, XT, YT>(s);
        }
    }

    void record(const BatchShape &sh, hipStream_t s) {
        drop_graph();
        build_items(sh.fields_per_item, s);
        shape = sh;
        using apply_detail::kAcc64;
        using apply_detail::kComp32;
        using apply_detail::kPlain;
        if (mode == ApplyMode::Plain) return record_xy<kPlain, double, double>(sh.fields_per_step, s);
        const bool bf = xt == FieldType::BF16, y64 = yt == FieldType::F64;
        if (mode == ApplyMode::Mixed) {
            if (bf) return y64 ? record_xy<kComp32, std::uint16_t, double>(sh.fields_per_step, s) : record_xy<kComp32, std::uint16_t, float>(sh.fields_per_step, s);
            return y64 ? record_xy<kComp32, float, double>(sh.fields_per_step, s) : record_xy<kComp32, float, float>(sh.fields_per_step, s);
        }
        if (bf) return y64 ? record_xy<kAcc64, std::uint16_t, double>(sh.fields_per_step, s) : record_xy<kAcc64, std::uint16_t, float>(sh.fields_per_step, s);
        return y64 ? record_xy<kAcc64, float, double>(sh.fields_per_step, s) : record_xy<kAcc64, float, float>(sh.fields_per_step, s);
    }

    // The pairs' own tuned launches, captured in order; each pair reads and
    // writes the plan's partial-sum workspace, as the persistent kernel does.
    std::vector<apply_detail::Launch> pair_shape;
    void record_pairs() {
        drop_graph();
        n_items = 0;
        n_wg = 0;
        AXIS_HIP_CHECK(hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal));
        for (std::size_t p = 0; p < mats.size(); ++p) {
            if (k[p] <= 0 || mats[p]->n_dst == 0) continue;
            apply_detail::SpecialArgs sa;
            sa.partials = partial[p].get();
            apply_detail::launch(*mats[p], src[p], xt, dst[p], yt, k[p], mode, pair_shape[p], cap, sa);
        }
        hipError_t launch_err = hipGetLastError();
        hipGraph_t g = nullptr;
        hipError_t cap_err = hipStreamEndCapture(cap, &g);
        AXIS_HIP_CHECK(launch_err);
        AXIS_HIP_CHECK(cap_err);
        graph = g;
        AXIS_HIP_CHECK(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
        shape = BatchShape{0, 0, true};
    }

    void launch(hipStream_t s) { AXIS_HIP_CHECK(hipGraphLaunch(exec, s)); }
};

BatchPlan::BatchPlan(std::vector<std::shared_ptr<InterpolationMatrix>> mats, std::vector<const void *> src, std::vector<void *> dst,
                     std::vector<index_t> k, FieldType xt, FieldType yt, ApplyMode mode, int autotune, hipStream_t s)
    : impl_(std::make_unique<Impl>()) {
    const std::size_t n = mats.size();
    if (n == 0) throw InvalidArgument("BatchPlan: no regrid pairs");
    if (src.size() != n || dst.size() != n || k.size() != n) throw InvalidArgument("BatchPlan: pair lists of different lengths");
    if (mode == ApplyMode::Plain && (xt != FieldType::F64 || yt != FieldType::F64))
        throw InvalidArgument("BatchPlan: the plain apply takes FP64 fields");
    if (mode != ApplyMode::Plain && (xt == FieldType::F64 || yt == FieldType::BF16))
        throw InvalidArgument("BatchPlan: the mixed apply takes FP32/BF16 fields and writes FP32/FP64");
    for (std::size_t p = 0; p < n; ++p) {
        if (!mats[p]) throw InvalidArgument("BatchPlan: null matrix");
        if (k[p] < 0 || k[p] >= (index_t(1) << 31)) throw InvalidArgument("BatchPlan: field count out of range");
        if (k[p] > 0 && (!src[p] || !dst[p])) throw InvalidArgument("BatchPlan: null field pointer");
    }
    auto &I = *impl_;
    I.mats = std::move(mats);
    I.src = std::move(src);
    I.dst = std::move(dst);
    I.k = std::move(k);
    I.xt = xt;
    I.yt = yt;
    I.mode = mode;
    I.d_ctr.resize(kQueues);
    AXIS_HIP_CHECK(hipStreamCreateWithFlags(&I.cap, hipStreamNonBlocking));
    I.partial.resize(n);
    std::vector<PairDesc> pd(n);
    index_t k_max = 0;
    for (std::size_t p = 0; p < n; ++p) {
        auto &A = *I.mats[p];
        {
            std::lock_guard<std::mutex> lock(A.apply_mu);
            A.prepare_tiles(s);
        }
        I.max_e = std::max(I.max_e, A.seg_cap);
        pd[p].A = apply_detail::mat_of(A);
        pd[p].X = I.src[p];
        pd[p].Y = I.dst[p];
        I.partial[p].resize(static_cast<std::size_t>(std::max(A.n_slots, 1)) * static_cast<std::size_t>(std::max<index_t>(I.k[p], 1)));
        pd[p].P = I.partial[p].get();
        k_max = std::max(k_max, I.k[p]);
    }
    I.d_pairs.resize(n);
    I.d_pairs.upload(pd.data(), n, s);
    AXIS_HIP_CHECK(hipStreamSynchronize(s));
    BatchShape best;
    best.fields_per_step = mode == ApplyMode::Mixed ? 4 : 2;
    best.fields_per_item = static_cast<int>(std::clamp<index_t>(k_max, 1, 16));
    // each pair's launch shape: its own autotuned (or cached, else heuristic) shape
    I.pair_shape.resize(n);
    for (std::size_t p = 0; p < n; ++p) {
        if (I.k[p] <= 0) continue;
        if (autotune != 0) {
            const ApplyShape sh = autotune_apply(*I.mats[p], I.src[p], xt, I.dst[p], yt, I.k[p], mode, s);
            I.pair_shape[p] = apply_detail::Launch{sh.fields_per_step, sh.field_chunks, sh.staged && !sh.wave, sh.wave};
        } else {
            I.pair_shape[p] = apply_detail::choose(*I.mats[p], I.k[p], xt, yt, mode);
        }
        const apply_detail::Launch &L = I.pair_shape[p];
        std::lock_guard<std::mutex> lock(I.mats[p]->apply_mu);
        if (apply_detail::runs_wave(*I.mats[p], L, xt, I.src[p]))
            I.mats[p]->prepare_wave_tiles(apply_detail::chunk_cols_of(xt), s);
        else if (L.tiled || L.wave)
            I.mats[p]->prepare_tiles(s);
    }
    AXIS_HIP_CHECK(hipStreamSynchronize(s));
    if (autotune == 1) {
        hipEvent_t e0, e1;
        AXIS_HIP_CHECK(hipEventCreate(&e0));
        AXIS_HIP_CHECK(hipEventCreate(&e1));
        float best_ms = 1e30f;
        auto time_graph = [&]() {
            I.launch(s);  // warm
            AXIS_HIP_CHECK(hipEventRecord(e0, s));
            for (int r = 0; r < 3; ++r) I.launch(s);
            AXIS_HIP_CHECK(hipEventRecord(e1, s));
            AXIS_HIP_CHECK(hipEventSynchronize(e1));
            float ms = 0.0f;
            AXIS_HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
            return ms;
        };
        for (int fu : {1, 2, 4})
            for (int fpi : {4, 8, 16, 32, 64, 128}) {
                if (fpi > 4 && fpi / 2 >= k_max) continue;
                if (fu > fpi) continue;
                I.record(BatchShape{fu, fpi}, s);
                const float ms = time_graph();
                if (std::getenv("AXIS_TRACE")) std::fprintf(stderr, "axis[trace] batch_plan.persistent fu %d fields/item %d: %.3f ms per 3 replays\n", fu, fpi, ms);
                if (ms < best_ms) {
                    best_ms = ms;
                    best = BatchShape{fu, fpi};
                }
            }
        I.record_pairs();
        const float ms = time_graph();
        if (std::getenv("AXIS_TRACE")) std::fprintf(stderr, "axis[trace] batch_plan.pair_launches: %.3f ms per 3 replays\n", ms);
        if (ms < best_ms) best = I.shape;
        AXIS_HIP_CHECK(hipEventDestroy(e0));
        AXIS_HIP_CHECK(hipEventDestroy(e1));
    }
    if (best.pair_launches || autotune == 2)
        I.record_pairs();
    else
        I.record(best, s);
}

BatchPlan::~BatchPlan() = default;

void BatchPlan::run(hipStream_t s, bool sync) {
    impl_->launch(s);
    if (sync) AXIS_HIP_CHECK(hipStreamSynchronize(s));
}

BatchShape BatchPlan::shape() const { return impl_->shape; }
index_t BatchPlan::n_items() const { return impl_->n_items; }
int BatchPlan::n_workgroups() const { return impl_->n_wg; }

}  // namespace axis_hip
