// ref_geometry_driver.cpp -- C-ABI shim over the reference's own
// header-only sphere geometry (libs/axis/include/axis/detail/spherical_geometry.hpp),
// compiled from the reference tree where it lies (never copied).
//
// TEST INFRASTRUCTURE ONLY: it exists to pin the oracle.
//   * ref_lonlat_to_xyz is the exact function extract_cell_polygon_spherical
//     calls on the weight-generation path (weight_generator.cpp:467), so it
//     pins the oracle's vertex conversion bit for bit.
//   * ref_polygon_area / ref_overlap_area are an independent clip/area
//     implementation (robust predicates, spherical_geometry.hpp:363-587);
//     they agree with the runtime clipper to ~1e-8 relative, not bitwise.
#include <axis/detail/spherical_geometry.hpp>
#include <vector>

using axis::detail::spherical::Vec3;

static std::vector<Vec3> load(const double *xyz, int n) {
    std::vector<Vec3> v(n);
    for (int i = 0; i < n; ++i) v[i] = Vec3{xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]};
    return v;
}

extern "C" {
void ref_lonlat_to_xyz(double lon_rad, double lat_rad, double *out) {
    Vec3 p = axis::detail::spherical::lonlat_to_xyz(lon_rad, lat_rad);
    out[0] = p.x;
    out[1] = p.y;
    out[2] = p.z;
}
double ref_polygon_area(const double *xyz, int n) { return axis::detail::spherical::spherical_polygon_area(load(xyz, n)); }
double ref_overlap_area(const double *subj, int ns, const double *clip, int nc) {
    return axis::detail::spherical::spherical_polygon_overlap_area(load(subj, ns), load(clip, nc));
}
}
