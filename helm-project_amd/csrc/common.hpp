// common.hpp -- HIP plumbing shared by every translation unit of libaxis_hip.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <cstring>
#include <memory>
#include <new>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

namespace axis_hip {

using index_t = std::int64_t;  // reference libs/axis/include/axis/types.hpp:90

/// Allocator whose resize() leaves new elements uninitialised, so that large
/// host arrays are first touched (page-faulted) by the threads that fill them
/// rather than zero-filled serially.
template <class T>
struct NoInitAlloc : std::allocator<T> {
    template <class U>
    struct rebind {
        using other = NoInitAlloc<U>;
    };
    NoInitAlloc() = default;
    template <class U>
    NoInitAlloc(const NoInitAlloc<U> &) noexcept {}
    template <class U, class... A>
    void construct(U *p, A &&...a) {
        if constexpr (sizeof...(A) == 0)
            ::new (static_cast<void *>(p)) U;
        else
            ::new (static_cast<void *>(p)) U(std::forward<A>(a)...);
    }
};
using HostDoubles = std::vector<double, NoInitAlloc<double>>;

/// Error raised for invalid arguments (maps to std::invalid_argument in the reference).
struct InvalidArgument : std::invalid_argument {
    using std::invalid_argument::invalid_argument;
};
/// Error raised for runtime failures: HIP errors, unmapped cells (std::runtime_error in the reference).
struct RuntimeError : std::runtime_error {
    using std::runtime_error::runtime_error;
};

[[noreturn]] inline void hip_fail(hipError_t e, const char *what, const char *file, int line) {
    throw RuntimeError(std::string("HIP error ") + hipGetErrorName(e) + " (" + hipGetErrorString(e) + ") in " + what + " at " + file + ":" +
                       std::to_string(line));
}

#define AXIS_HIP_CHECK(expr)                                                      \
    do {                                                                          \
        hipError_t _e = (expr);                                                   \
        if (_e != hipSuccess) ::axis_hip::hip_fail(_e, #expr, __FILE__, __LINE__); \
    } while (0)

#define AXIS_LAUNCH_CHECK() AXIS_HIP_CHECK(hipGetLastError())

/// Process-wide cache of device blocks.  hipFree unmaps memory and waits for
/// the device, which costs ~0.4 ms per block; a generation or a matrix's
/// lifetime creates dozens.  Released blocks are parked as pending; the first
/// allocation that needs one waits for the device once and recycles all of
/// them, so a block is never reused while a kernel queued before its release
/// may still touch it.  Blocks are sized in steps (powers of two up to 2 MiB,
/// then multiples of 2 MiB) and reused for requests at least half their size.
class BlockCache {
   public:
    static BlockCache &get();
    void *alloc(std::size_t bytes);
    void release(void *p, std::size_t bytes);
    /// Returns every cached block to the runtime (waits for the device).
    void trim();
    /// While one of these lives on a thread, that thread's allocations never
    /// wait for the device to recycle released blocks (a miss is served by a
    /// new block instead): for pipelines whose other streams (host-fed
    /// uploads) must not be waited for.  Device and pinned blocks alike.
    struct NoWaitScope {
        NoWaitScope();
        ~NoWaitScope();
        NoWaitScope(const NoWaitScope &) = delete;
        NoWaitScope &operator=(const NoWaitScope &) = delete;
    };
    static bool no_wait();
    /// Waits for the device once and makes every released device and pinned
    /// block reusable: called where the device is about to be idle anyway,
    /// before a NoWaitScope, so that the scope's allocations find blocks.
    static void recycle_now();
    static std::size_t rounded(std::size_t bytes) {
        if (bytes <= (std::size_t(2) << 20)) {
            std::size_t r = 256;
            while (r < bytes) r <<= 1;
            return r;
        }
        const std::size_t step = std::size_t(2) << 20;
        return (bytes + step - 1) / step * step;
    }
};

/// Owning device allocation from the block cache, movable, not copyable.
/// Process-wide cache of pinned host blocks (hipHostMalloc pins pages, which
/// costs milliseconds for the tens of MB a mesh's node coordinates take).
/// Same policy as BlockCache: released blocks wait for one device
/// synchronisation before reuse, so an upload queued from a block has finished
/// when the block is handed out again.
class PinnedCache {
   public:
    static PinnedCache &get();
    void *alloc(std::size_t bytes);
    void release(void *p, std::size_t bytes);
};

/// Uninitialised pinned host array (resize discards the contents).
template <class T>
class PinnedArray {
   public:
    PinnedArray() = default;
    ~PinnedArray() { clear(); }
    PinnedArray(const PinnedArray &) = delete;
    PinnedArray &operator=(const PinnedArray &) = delete;
    PinnedArray(PinnedArray &&o) noexcept : p_(std::exchange(o.p_, nullptr)), n_(std::exchange(o.n_, 0)) {}
    PinnedArray &operator=(PinnedArray &&o) noexcept {
        if (this != &o) {
            clear();
            p_ = std::exchange(o.p_, nullptr);
            n_ = std::exchange(o.n_, 0);
        }
        return *this;
    }
    void resize(std::size_t n) {
        clear();
        if (n) p_ = static_cast<T *>(PinnedCache::get().alloc(n * sizeof(T)));
        n_ = n;
    }
    void clear() {
        if (p_) PinnedCache::get().release(p_, n_ * sizeof(T));
        p_ = nullptr;
        n_ = 0;
    }
    T *data() { return p_; }
    const T *data() const { return p_; }
    T &operator[](std::size_t i) { return p_[i]; }
    const T &operator[](std::size_t i) const { return p_[i]; }
    std::size_t size() const { return n_; }
    bool empty() const { return n_ == 0; }

   private:
    T *p_ = nullptr;
    std::size_t n_ = 0;
};

template <class T>
class DeviceBuffer {
   public:
    DeviceBuffer() = default;
    explicit DeviceBuffer(std::size_t n) { resize(n); }
    ~DeviceBuffer() { reset(); }
    DeviceBuffer(const DeviceBuffer &) = delete;
    DeviceBuffer &operator=(const DeviceBuffer &) = delete;
    DeviceBuffer(DeviceBuffer &&o) noexcept : p_(std::exchange(o.p_, nullptr)), n_(std::exchange(o.n_, 0)) {}
    DeviceBuffer &operator=(DeviceBuffer &&o) noexcept {
        if (this != &o) {
            reset();
            p_ = std::exchange(o.p_, nullptr);
            n_ = std::exchange(o.n_, 0);
        }
        return *this;
    }
    void resize(std::size_t n) {
        reset();
        if (n) p_ = static_cast<T *>(BlockCache::get().alloc(n * sizeof(T)));
        n_ = n;
    }
    void reset() {
        if (p_) BlockCache::get().release(p_, n_ * sizeof(T));
        p_ = nullptr;
        n_ = 0;
    }
    T *get() const { return p_; }
    std::size_t size() const { return n_; }
    std::size_t bytes() const { return n_ * sizeof(T); }
    void upload(const T *h, std::size_t n, hipStream_t s) {
        if (n > n_) throw InvalidArgument("DeviceBuffer::upload overflow");
        if (n) AXIS_HIP_CHECK(hipMemcpyAsync(p_, h, n * sizeof(T), hipMemcpyHostToDevice, s));
    }
    void download(T *h, std::size_t n, hipStream_t s) const {
        if (n > n_) throw InvalidArgument("DeviceBuffer::download overflow");
        if (n) AXIS_HIP_CHECK(hipMemcpyAsync(h, p_, n * sizeof(T), hipMemcpyDeviceToHost, s));
    }

   private:
    T *p_ = nullptr;
    std::size_t n_ = 0;
};

/// Copies n Ts from device memory at d to h (host memory of any kind) and
/// waits for `s`.  Copies of up to 4 KiB go through a per-thread pinned
/// staging buffer: a pageable destination makes the runtime stage the copy on
/// its own threads, which a busy host delays.
template <class T>
inline void read_back(T *h, const T *d, std::size_t n, hipStream_t s) {
    constexpr std::size_t kStage = 4096;
    static thread_local void *stage = nullptr;  // (one per thread, kept until the process ends)
    const std::size_t bytes = n * sizeof(T);
    if (bytes == 0) return;
    if (bytes > kStage) {
        AXIS_HIP_CHECK(hipMemcpyAsync(h, d, bytes, hipMemcpyDeviceToHost, s));
        AXIS_HIP_CHECK(hipStreamSynchronize(s));
        return;
    }
    if (!stage) AXIS_HIP_CHECK(hipHostMalloc(&stage, kStage, hipHostMallocDefault));
    AXIS_HIP_CHECK(hipMemcpyAsync(stage, d, bytes, hipMemcpyDeviceToHost, s));
    AXIS_HIP_CHECK(hipStreamSynchronize(s));
    std::memcpy(h, stage, bytes);
}

inline unsigned grid_for(std::size_t n, unsigned block) {
    std::size_t g = (n + block - 1) / block;
    return static_cast<unsigned>(g ? g : 1);
}

}  // namespace axis_hip
