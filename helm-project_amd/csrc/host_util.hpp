// host_util.hpp -- small host-side helpers (thread pool loop, libm-exact
// coordinate conversion shared by every host step that must reproduce the
// reference's bits).
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <atomic>
#include <condition_variable>
#include <deque>
#include <exception>
#include <functional>
#include <mutex>
#include <thread>
#include <vector>

#include "geometry.hpp"

namespace axis_hip {

/// Host threads used for the libm-exact O(n) steps: AXIS_HOST_THREADS, else
/// hardware_concurrency capped at 16.
inline int host_threads() {
    static int n = [] {
        if (const char *e = std::getenv("AXIS_HOST_THREADS")) {
            int v = std::atoi(e);
            if (v > 0) return v;
        }
        unsigned hc = std::thread::hardware_concurrency();
        return static_cast<int>(std::clamp(hc ? hc : 1u, 1u, 16u));
    }();
    return n;
}

namespace pool_detail {

/// One parallel_range call: chunk indices handed out by an atomic counter.
struct Job {
    std::function<void(std::int64_t)> fn;
    std::int64_t chunks = 0;
    std::atomic<std::int64_t> next{0};
    std::int64_t done = 0;  // guarded by mu
    int refs = 0;           // workers inside the job, guarded by the pool's mutex
    std::exception_ptr err; // guarded by mu
    std::mutex mu;
    std::condition_variable cv;
};

/// Persistent host workers (created once, host_threads() - 1 of them; the
/// caller of run() works on its own job too, so a call always makes progress
/// even when every worker is busy with another caller's job).
class Pool {
   public:
    explicit Pool(int workers) {
        for (int i = 0; i < workers; ++i) std::thread([this] { loop(); }).detach();
        workers_ = workers;
    }
    void run(Job &j) {
        {
            std::lock_guard<std::mutex> g(mu_);
            const std::int64_t helpers = std::min<std::int64_t>(workers_, j.chunks - 1);
            for (std::int64_t i = 0; i < helpers; ++i) q_.push_back(&j);
        }
        cv_.notify_all();
        work(j);
        {
            std::unique_lock<std::mutex> l(j.mu);
            j.cv.wait(l, [&] { return j.done == j.chunks; });
        }
        {
            // queued entries nobody took are dropped; wait for workers still inside
            std::unique_lock<std::mutex> l(mu_);
            q_.erase(std::remove(q_.begin(), q_.end(), &j), q_.end());
            left_.wait(l, [&] { return j.refs == 0; });
        }
        if (j.err) std::rethrow_exception(j.err);
    }

   private:
    static void work(Job &j) {
        for (;;) {
            const std::int64_t c = j.next.fetch_add(1);
            if (c >= j.chunks) return;
            try {
                j.fn(c);
            } catch (...) {
                std::lock_guard<std::mutex> g(j.mu);
                if (!j.err) j.err = std::current_exception();
            }
            std::lock_guard<std::mutex> g(j.mu);
            if (++j.done == j.chunks) j.cv.notify_all();
        }
    }
    void loop() {
        for (;;) {
            Job *j = nullptr;
            {
                std::unique_lock<std::mutex> l(mu_);
                cv_.wait(l, [&] { return !q_.empty(); });
                j = q_.front();
                q_.pop_front();
                ++j->refs;
            }
            work(*j);
            {
                std::lock_guard<std::mutex> l(mu_);
                --j->refs;
            }
            left_.notify_all();
        }
    }
    std::mutex mu_;
    std::condition_variable cv_, left_;
    std::deque<Job *> q_;
    int workers_ = 0;
};

inline Pool &pool() {
    static Pool *p = new Pool(std::max(host_threads() - 1, 0));  // never destroyed: the workers outlive static destruction
    return *p;
}

}  // namespace pool_detail

/// Runs fn(begin, end) over [0, n) split into contiguous chunks (at most
/// max_threads of them when given), on the persistent host workers and the
/// calling thread.  An exception thrown by fn is rethrown here after every
/// chunk has finished.
template <class Fn>
void parallel_range(std::int64_t n, Fn fn, std::int64_t min_chunk = 4096, int max_threads = 0) {
    int nt = max_threads > 0 ? std::min(max_threads, host_threads()) : host_threads();
    std::int64_t chunks = std::max<std::int64_t>(1, std::min<std::int64_t>(nt, (n + min_chunk - 1) / min_chunk));
    if (chunks <= 1) {
        fn(std::int64_t(0), n);
        return;
    }
    pool_detail::Job j;
    j.chunks = chunks;
    j.fn = [&](std::int64_t c) { fn(n * c / chunks, n * (c + 1) / chunks); };
    pool_detail::pool().run(j);
}

/// Phase timer for AXIS_TRACE=1: each mark() waits for the stream and prints
/// the wall time since the previous mark to stderr (the role Kokkos Tools
/// region hooks play for the reference).  Disabled, it costs one branch.
class PhaseTrace {
   public:
    PhaseTrace(const char *scope, hipStream_t s) : scope_(scope), s_(s), on_(std::getenv("AXIS_TRACE") != nullptr) {
        if (on_) last_ = std::chrono::steady_clock::now();
    }
    void mark(const char *what) {
        if (!on_) return;
        (void)hipStreamSynchronize(s_);
        auto now = std::chrono::steady_clock::now();
        std::fprintf(stderr, "axis[trace] %s.%s %.3f ms\n", scope_, what, std::chrono::duration<double, std::milli>(now - last_).count());
        last_ = now;
    }

   private:
    const char *scope_;
    hipStream_t s_;
    bool on_;
    std::chrono::steady_clock::time_point last_;
};

constexpr double kPi = 3.14159265358979323846;
constexpr double kDeg2Rad = kPi / 180.0;  // folded once, as spherical_geometry.hpp:36 / spherical_cap_filter.hpp:40

/// sin and cos of one argument through glibc's sincos.  The reference's
/// sin/cos pairs (lonlat_to_xyz, the trig cache, the cap filter) are emitted by
/// GCC at -O2 and above as one sincos call, and glibc's sincos differs from
/// separate sin/cos in the last bit for ~0.1% of arguments; those bits decide
/// sliver overlaps, so every host pair goes through sincos too.
inline void sin_cos(double x, double &s, double &c) { ::sincos(x, &s, &c); }

/// lonlat_to_xyz (spherical_geometry.hpp:80-83) with host libm.
inline geo::V3 lonlat_to_xyz_host(double lon_rad, double lat_rad) {
    double sl, cl, so, co;
    sin_cos(lat_rad, sl, cl);
    sin_cos(lon_rad, so, co);
    return geo::V3{cl * co, cl * so, sl};
}

/// Direct-mapped memo of sin_cos keyed by the argument's bits (the same bits
/// in give the same bits out), for node coordinates that repeat: every row of
/// a lat-lon grid shares one latitude and every column one longitude.
class SinCosMemo {
   public:
    SinCosMemo() : e_(kSize) {
        for (auto &x : e_) x.key = kEmpty;
    }
    void get(double x, double &s, double &c) {
        std::uint64_t b;
        std::memcpy(&b, &x, sizeof b);
        if (b == kEmpty) {
            sin_cos(x, s, c);
            return;
        }
        Entry &e = e_[(b * 0x9E3779B97F4A7C15ull) >> (64 - kBits)];
        if (e.key != b) {
            sin_cos(x, e.s, e.c);
            e.key = b;
        }
        s = e.s;
        c = e.c;
    }
    /// lonlat_to_xyz_host through the memo
    geo::V3 xyz(double lon_rad, double lat_rad) {
        double sl, cl, so, co;
        get(lat_rad, sl, cl);
        get(lon_rad, so, co);
        return geo::V3{cl * co, cl * so, sl};
    }

   private:
    static constexpr int kBits = 12;
    static constexpr std::size_t kSize = std::size_t(1) << kBits;
    static constexpr std::uint64_t kEmpty = ~std::uint64_t(0);  // a NaN pattern, passed straight through
    struct Entry {
        std::uint64_t key;
        double s, c;
    };
    std::vector<Entry> e_;
};

}  // namespace axis_hip
