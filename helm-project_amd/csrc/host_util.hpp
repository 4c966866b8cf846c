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

/// Runs fn(begin, end) over [0, n) split into contiguous chunks.
template <class Fn>
void parallel_range(std::int64_t n, Fn fn, std::int64_t min_chunk = 4096) {
    int nt = host_threads();
    std::int64_t chunks = std::max<std::int64_t>(1, std::min<std::int64_t>(nt, (n + min_chunk - 1) / min_chunk));
    if (chunks <= 1) {
        fn(std::int64_t(0), n);
        return;
    }
    std::vector<std::thread> th;
    th.reserve(chunks);
    for (std::int64_t c = 0; c < chunks; ++c) {
        std::int64_t b = n * c / chunks, e = n * (c + 1) / chunks;
        th.emplace_back([=] { fn(b, e); });
    }
    for (auto &t : th) t.join();
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
