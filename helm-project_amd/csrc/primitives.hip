// primitives.hip -- device-wide exclusive scan and LSD radix sort (gfx950,
// 64-wide wavefronts).  Used by the LBVH (Morton sort), the CSR assembly
// (row counts -> row_ptr) and the grid detector (sorted node coordinates).
#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>
#include <mutex>
#include <map>
#include <unordered_map>

#include "axis_hip.hpp"

namespace axis_hip {

// ------------------------------------------------------------- block cache
namespace {
struct Block {
    int device;
    std::size_t bytes;
    void *p;
};
struct CacheState {
    std::mutex mu;
    std::multimap<std::pair<int, std::size_t>, void *> free;  // (device, rounded size) -> block
    std::vector<Block> pending;                               // released, possibly still in use by queued work
    std::unordered_map<void *, int> owner;                    // live block -> device
    void recycle() {
        for (const Block &b : pending) free.emplace(std::make_pair(b.device, b.bytes), b.p);
        pending.clear();
    }
};
int current_device() {
    int d = 0;
    AXIS_HIP_CHECK(hipGetDevice(&d));
    return d;
}
CacheState &cache_state() {
    static CacheState *c = new CacheState;  // never destroyed: blocks live until the process ends
    return *c;
}
CacheState &pinned_state() {
    static CacheState *c = new CacheState;
    return *c;
}
}  // namespace

BlockCache &BlockCache::get() {
    static BlockCache b;
    return b;
}

namespace {
thread_local int g_no_wait = 0;
}
BlockCache::NoWaitScope::NoWaitScope() { ++g_no_wait; }
BlockCache::NoWaitScope::~NoWaitScope() { --g_no_wait; }
bool BlockCache::no_wait() { return g_no_wait > 0; }

void BlockCache::recycle_now() {
    CacheState &c = cache_state(), &h = pinned_state();
    std::scoped_lock g(c.mu, h.mu);
    if (c.pending.empty() && h.pending.empty()) return;
    AXIS_HIP_CHECK(hipDeviceSynchronize());
    c.recycle();
    h.recycle();
}

void *BlockCache::alloc(std::size_t bytes) {
    const std::size_t r = rounded(bytes);
    const int dev = current_device();
    CacheState &c = cache_state();
    std::lock_guard<std::mutex> g(c.mu);
    for (int pass = 0; pass < 2; ++pass) {
        auto it = c.free.lower_bound(std::make_pair(dev, r));
        if (it != c.free.end() && it->first.first == dev && it->first.second <= 2 * r) {
            void *p = it->second;
            c.free.erase(it);
            c.owner[p] = dev;
            return p;
        }
        if (pass == 0 && !c.pending.empty() && !no_wait()) {
            AXIS_HIP_CHECK(hipDeviceSynchronize());  // work queued before the releases has finished
            c.recycle();
            continue;
        }
        break;
    }
    void *p = nullptr;
    hipError_t e = hipMalloc(&p, r);
    if (e != hipSuccess) {  // out of memory: give this device's cached blocks back and retry once
        (void)hipGetLastError();
        AXIS_HIP_CHECK(hipDeviceSynchronize());
        c.recycle();
        for (auto it = c.free.begin(); it != c.free.end();) {
            if (it->first.first == dev) {
                (void)hipFree(it->second);
                it = c.free.erase(it);
            } else {
                ++it;
            }
        }
        AXIS_HIP_CHECK(hipMalloc(&p, r));
    }
    c.owner[p] = dev;
    return p;
}

void BlockCache::release(void *p, std::size_t bytes) {
    CacheState &c = cache_state();
    std::lock_guard<std::mutex> g(c.mu);
    int dev = 0;
    if (auto it = c.owner.find(p); it != c.owner.end()) {
        dev = it->second;
        c.owner.erase(it);
    }
    c.pending.push_back(Block{dev, rounded(bytes), p});
}

PinnedCache &PinnedCache::get() {
    static PinnedCache b;
    return b;
}

void *PinnedCache::alloc(std::size_t bytes) {
    const std::size_t r = BlockCache::rounded(bytes);
    CacheState &c = pinned_state();
    std::lock_guard<std::mutex> g(c.mu);
    for (int pass = 0; pass < 2; ++pass) {
        auto it = c.free.lower_bound(std::make_pair(0, r));
        if (it != c.free.end() && it->first.second <= 2 * r) {
            void *p = it->second;
            c.free.erase(it);
            return p;
        }
        if (pass == 0 && !c.pending.empty() && !BlockCache::no_wait()) {
            AXIS_HIP_CHECK(hipDeviceSynchronize());  // uploads queued from the released blocks have finished
            c.recycle();
            continue;
        }
        break;
    }
    void *p = nullptr;
    AXIS_HIP_CHECK(hipHostMalloc(&p, r, hipHostMallocDefault));
    return p;
}

void PinnedCache::release(void *p, std::size_t bytes) {
    CacheState &c = pinned_state();
    std::lock_guard<std::mutex> g(c.mu);
    c.pending.push_back(Block{0, BlockCache::rounded(bytes), p});
}

void BlockCache::trim() {
    CacheState &c = cache_state();
    std::lock_guard<std::mutex> g(c.mu);
    AXIS_HIP_CHECK(hipDeviceSynchronize());
    c.recycle();
    for (auto &b : c.free) (void)hipFree(b.second);
    c.free.clear();
}

}  // namespace axis_hip

namespace axis_hip::detail {

namespace {

constexpr int kScanBlock = 256;
constexpr int kScanItems = 8;
constexpr int kScanTile = kScanBlock * kScanItems;

// Block-wide inclusive scan of one value per thread through LDS (wave scans
// with DPP-backed shuffles, then a scan of the four wave totals).
__device__ inline index_t block_inclusive_scan(index_t v, index_t *lds_waves, index_t &block_total) {
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        index_t t = __shfl_up(v, off, 64);
        if (lane >= off) v += t;
    }
    if (lane == 63) lds_waves[wave] = v;
    __syncthreads();
    index_t prefix = 0, total = 0;
#pragma unroll
    for (int w = 0; w < kScanBlock / 64; ++w) {
        index_t s = lds_waves[w];
        if (w < wave) prefix += s;
        total += s;
    }
    __syncthreads();
    block_total = total;
    return v + prefix;
}

__global__ void __launch_bounds__(kScanBlock) scan_reduce_kernel(const index_t *in, index_t n, index_t *block_sums) {
    __shared__ index_t lds[kScanBlock / 64];
    const index_t base = static_cast<index_t>(blockIdx.x) * kScanTile;
    index_t s = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        index_t i = base + static_cast<index_t>(k) * kScanBlock + threadIdx.x;
        if (i < n) s += in[i];
    }
    index_t total;
    block_inclusive_scan(s, lds, total);
    if (threadIdx.x == 0) block_sums[blockIdx.x] = total;
}

// Each thread owns kScanItems consecutive elements.
__global__ void __launch_bounds__(kScanBlock) scan_down_kernel(const index_t *in, index_t n, const index_t *block_offsets, index_t *out) {
    __shared__ index_t lds[kScanBlock / 64];
    const index_t base = static_cast<index_t>(blockIdx.x) * kScanTile + static_cast<index_t>(threadIdx.x) * kScanItems;
    index_t v[kScanItems];
    index_t s = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        index_t i = base + k;
        v[k] = (i < n) ? in[i] : 0;
        s += v[k];
    }
    index_t total;
    index_t incl = block_inclusive_scan(s, lds, total);
    index_t run = block_offsets[blockIdx.x] + incl - s;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        index_t i = base + k;
        if (i < n) out[i] = run;
        run += v[k];
    }
    if (blockIdx.x == gridDim.x - 1 && threadIdx.x == kScanBlock - 1) out[n] = run;
}

}  // namespace

namespace {
// grid-stride copies; a host pointer is a mapped pinned block
__global__ void copy_u64_kernel(std::uint64_t *dst, const std::uint64_t *src, std::size_t n) {
    for (std::size_t i = static_cast<std::size_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += static_cast<std::size_t>(gridDim.x) * blockDim.x)
        dst[i] = src[i];
}
__global__ void copy_u8_kernel(unsigned char *dst, const unsigned char *src, std::size_t n) {
    for (std::size_t i = static_cast<std::size_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += static_cast<std::size_t>(gridDim.x) * blockDim.x)
        dst[i] = src[i];
}
}  // namespace

void kernel_copy(void *dst, const void *src, std::size_t bytes, hipStream_t s) {
    if (bytes == 0) return;
    const bool words = (reinterpret_cast<std::uintptr_t>(dst) % 8) == 0 && (reinterpret_cast<std::uintptr_t>(src) % 8) == 0 && bytes % 8 == 0;
    const std::size_t n = words ? bytes / 8 : bytes;
    const unsigned grid = static_cast<unsigned>(std::min<std::size_t>((n + 255) / 256, 1024));
    if (words)
        copy_u64_kernel<<<grid, 256, 0, s>>>(static_cast<std::uint64_t *>(dst), static_cast<const std::uint64_t *>(src), n);
    else
        copy_u8_kernel<<<grid, 256, 0, s>>>(static_cast<unsigned char *>(dst), static_cast<const unsigned char *>(src), n);
    AXIS_LAUNCH_CHECK();
}

void exclusive_scan_i64(const index_t *in, index_t *out, index_t n, hipStream_t s) {
    if (n <= 0) {
        AXIS_HIP_CHECK(hipMemsetAsync(out, 0, sizeof(index_t), s));
        return;
    }
    const index_t nb = (n + kScanTile - 1) / kScanTile;
    DeviceBuffer<index_t> sums(nb), offs(nb + 1);
    scan_reduce_kernel<<<static_cast<unsigned>(nb), kScanBlock, 0, s>>>(in, n, sums.get());
    AXIS_LAUNCH_CHECK();
    if (nb == 1) {
        AXIS_HIP_CHECK(hipMemsetAsync(offs.get(), 0, sizeof(index_t), s));
    } else {
        exclusive_scan_i64(sums.get(), offs.get(), nb, s);
    }
    scan_down_kernel<<<static_cast<unsigned>(nb), kScanBlock, 0, s>>>(in, n, offs.get(), out);
    AXIS_LAUNCH_CHECK();
    // (no wait: the temporaries go back to the block cache, which hands a
    // block out again only after a device synchronisation)
}

// --------------------------------------------------------------- radix sort
namespace {

constexpr int kSortBlock = 256;
constexpr int kSortWaves = kSortBlock / 64;
constexpr int kSortRounds = 16;
constexpr int kSortTile = kSortBlock * kSortRounds;

template <class K>
__global__ void __launch_bounds__(kSortBlock) radix_hist_kernel(const K *keys, std::size_t n, int shift, index_t *hist, unsigned nblocks) {
    __shared__ unsigned cnt[256];
    cnt[threadIdx.x] = 0;
    __syncthreads();
    const std::size_t base = static_cast<std::size_t>(blockIdx.x) * kSortTile;
    for (int r = 0; r < kSortRounds; ++r) {
        std::size_t i = base + static_cast<std::size_t>(r) * kSortBlock + threadIdx.x;
        if (i < n) atomicAdd(&cnt[static_cast<unsigned>((keys[i] >> shift) & 0xFF)], 1u);
    }
    __syncthreads();
    hist[static_cast<std::size_t>(threadIdx.x) * nblocks + blockIdx.x] = cnt[threadIdx.x];
}

// Stable scatter: keys are taken in index order, 256 per round; a key's rank
// among equal digits inside its wave comes from 8 ballots (one per digit bit),
// across waves from per-wave digit counts in LDS.
template <class K, bool kVals>
__global__ void __launch_bounds__(kSortBlock) radix_scatter_kernel(const K *keys_in, const std::uint32_t *vals_in, K *keys_out,
                                                                   std::uint32_t *vals_out, std::size_t n, int shift, const index_t *offsets,
                                                                   unsigned nblocks) {
    __shared__ unsigned run[256];
    __shared__ unsigned wcnt[kSortWaves][256];
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    run[threadIdx.x] = static_cast<unsigned>(offsets[static_cast<std::size_t>(threadIdx.x) * nblocks + blockIdx.x]);
    const std::size_t base = static_cast<std::size_t>(blockIdx.x) * kSortTile;
    const unsigned long long lt = (lane == 0) ? 0ull : ((~0ull) >> (64 - lane));
    for (int r = 0; r < kSortRounds; ++r) {
        for (int w = 0; w < kSortWaves; ++w) wcnt[w][threadIdx.x] = 0;
        __syncthreads();
        const std::size_t i = base + static_cast<std::size_t>(r) * kSortBlock + threadIdx.x;
        const bool valid = i < n;
        K key = valid ? keys_in[i] : K(0);
        const unsigned digit = static_cast<unsigned>((key >> shift) & 0xFF);
        unsigned long long peers = __ballot(valid);
#pragma unroll
        for (int b = 0; b < 8; ++b) {
            unsigned long long bal = __ballot((digit >> b) & 1u);
            peers &= ((digit >> b) & 1u) ? bal : ~bal;
        }
        const unsigned rank = __popcll(peers & lt);
        if (valid && rank == 0) wcnt[wave][digit] = __popcll(peers);
        __syncthreads();
        if (valid) {
            unsigned before = 0;
            for (int w = 0; w < wave; ++w) before += wcnt[w][digit];
            const unsigned pos = run[digit] + before + rank;
            keys_out[pos] = key;
            if constexpr (kVals) vals_out[pos] = vals_in[i];
        }
        __syncthreads();
        unsigned add = 0;
        for (int w = 0; w < kSortWaves; ++w) add += wcnt[w][threadIdx.x];
        run[threadIdx.x] += add;
        __syncthreads();
    }
}

// LSD passes over key bits [first_bit, bits); stable, so keys whose lower
// bits are already in order need only the upper passes
template <class K, bool kVals>
void radix_sort_impl(K *keys, std::uint32_t *vals, std::size_t n, int bits, hipStream_t s, int first_bit = 0) {
    if (n <= 1) return;
    if (n >= (std::size_t(1) << 32)) throw InvalidArgument("radix sort: more than 2^32 keys");
    const unsigned nb = static_cast<unsigned>((n + kSortTile - 1) / kSortTile);
    DeviceBuffer<K> k2(n);
    DeviceBuffer<std::uint32_t> v2(kVals ? n : 0);
    DeviceBuffer<index_t> hist(static_cast<std::size_t>(256) * nb), offs(static_cast<std::size_t>(256) * nb + 1);
    K *kin = keys, *kout = k2.get();
    std::uint32_t *vin = vals, *vout = v2.get();
    const int p0 = first_bit / 8;
    const int passes = (bits + 7) / 8;
    for (int p = p0; p < passes; ++p) {
        const int shift = 8 * p;
        radix_hist_kernel<K><<<nb, kSortBlock, 0, s>>>(kin, n, shift, hist.get(), nb);
        AXIS_LAUNCH_CHECK();
        exclusive_scan_i64(hist.get(), offs.get(), static_cast<index_t>(256) * nb, s);
        radix_scatter_kernel<K, kVals><<<nb, kSortBlock, 0, s>>>(kin, vin, kout, vout, n, shift, offs.get(), nb);
        AXIS_LAUNCH_CHECK();
        std::swap(kin, kout);
        std::swap(vin, vout);
    }
    if (kin != keys) {
        AXIS_HIP_CHECK(hipMemcpyAsync(keys, kin, n * sizeof(K), hipMemcpyDeviceToDevice, s));
        if constexpr (kVals) AXIS_HIP_CHECK(hipMemcpyAsync(vals, vin, n * sizeof(std::uint32_t), hipMemcpyDeviceToDevice, s));
    }
}

}  // namespace

void radix_sort_pairs_u32(std::uint32_t *keys, std::uint32_t *vals, std::size_t n, int bits, hipStream_t s) {
    radix_sort_impl<std::uint32_t, true>(keys, vals, n, bits, s);
}
void radix_sort_keys_u64(std::uint64_t *keys, std::size_t n, hipStream_t s, int bits, int first_bit) {
    radix_sort_impl<std::uint64_t, false>(keys, nullptr, n, bits, s, first_bit);
}
void radix_sort_pairs_u64(std::uint64_t *keys, std::uint32_t *vals, std::size_t n, int bits, hipStream_t s) {
    radix_sort_impl<std::uint64_t, true>(keys, vals, n, bits, s);
}

}  // namespace axis_hip::detail
