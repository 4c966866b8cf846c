// Host check of the double-double sin/cos used for node xyz on the device
// (helm-project_amd/csrc/ddtrig.hpp, compiled here for the CPU).
//
// Reads doubles (degrees) from argv[1]; for each x = deg * (pi / 180):
//  * hi + lo agrees with long double sinl/cosl to 2^-60 relative (the
//    double-double value is ~100 bits; sinl is within one 64-bit ULP);
//  * glibc's sincos returns hi whenever hi + lo lies at least kLibmBand ULP
//    from a rounding midpoint, and otherwise the double that pick() selects
//    from the last bit of glibc's answer (results below kTinyResult, left to
//    host libm, are counted only).
// Prints counts as one JSON line; exit status 1 on any failure.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <vector>

#include "../helm-project_amd/csrc/ddtrig.hpp"

using axis_hip::detail::dd::DD;
namespace dd = axis_hip::detail::dd;

int main(int argc, char **argv) {
    if (argc < 2) return 2;
    std::FILE *f = std::fopen(argv[1], "rb");
    if (!f) return 2;
    std::vector<double> v;
    double buf[4096];
    std::size_t got;
    while ((got = std::fread(buf, sizeof(double), 4096, f)) > 0) v.insert(v.end(), buf, buf + got);
    std::fclose(f);
    const double k = 3.14159265358979323846 / 180.0;
    long n_vals = 0, n_tiny = 0, n_out_of_range = 0, n_flagged = 0, n_libm_other = 0, bad_accuracy = 0, bad_libm = 0, bad_range = 0;
    double max_rel = 0.0, max_dist_other = 0.0;
    for (double deg : v) {
        const double x = deg * k;
        DD s, c;
        if (!dd::sincos_dd(x, s, c)) {
            ++n_out_of_range;
            if (std::fabs(x) <= dd::kMaxArg) ++bad_range;
            continue;
        }
        double gs, gc;
        ::sincos(x, &gs, &gc);
        const long double xl = x;
        const long double ref[2] = {sinl(xl), cosl(xl)};
        const DD got2[2] = {s, c};
        const double libm[2] = {gs, gc};
        for (int q = 0; q < 2; ++q) {
            ++n_vals;
            const long double val = static_cast<long double>(got2[q].hi) + static_cast<long double>(got2[q].lo);
            const long double den = std::fabs(ref[q]) > 0.0L ? std::fabs(ref[q]) : 1.0L;
            const double rel = static_cast<double>(std::fabs(val - ref[q]) / den);
            if (rel > max_rel) max_rel = rel;
            if (rel > std::ldexp(1.0, -60)) ++bad_accuracy;
            const double dist = dd::midpoint_distance(got2[q]);
            std::uint64_t lb;
            std::memcpy(&lb, &libm[q], sizeof lb);
            if (dd::tiny(got2[q])) {  // left to host libm
                ++n_tiny;
                continue;
            }
            if (!dd::near_midpoint(got2[q])) {
                if (libm[q] != got2[q].hi) ++bad_libm;
            } else {
                ++n_flagged;
                if (dd::pick(got2[q], static_cast<unsigned>(lb & 1u)) != libm[q]) ++bad_libm;
            }
            if (libm[q] != got2[q].hi) {
                ++n_libm_other;
                if (dist > max_dist_other) max_dist_other = dist;
            }
        }
    }
    std::printf(
        "{\"inputs\": %zu, \"values\": %ld, \"tiny\": %ld, \"out_of_range\": %ld, \"flagged\": %ld, \"libm_not_nearest\": %ld, "
        "\"max_midpoint_distance_libm_not_nearest\": %.6f, \"max_rel_err\": %.3e, \"bad_accuracy\": %ld, \"bad_libm\": %ld, \"bad_range\": %ld}\n",
        v.size(), n_vals, n_tiny, n_out_of_range, n_flagged, n_libm_other, max_dist_other, max_rel, bad_accuracy, bad_libm, bad_range);
    return (bad_accuracy || bad_libm || bad_range) ? 1 : 0;
}
