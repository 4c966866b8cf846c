// grid_descriptor.hpp -- host-side mirror of the reference's GridDescriptor
// (libs/axis/include/axis/ingest/grid_descriptor.hpp:37-216), the plain-data
// ingest contract whose address axis_mesh_from_descriptor_c receives as an
// opaque pointer (axis_c_interop.cpp:105-109).
//
// The caller's struct is read in place, so this mirror keeps the reference's
// member order and types: the enums are uint8_t, the strings are libstdc++
// std::string (the same C++ standard library on both sides of the link), and a
// field_view (types.hpp:69-70, Kokkos::mdspan with layout_left and the default
// accessor) is its data handle followed by one size_t per dynamic extent.  The
// offsets asserted below are the x86-64 System V layout of that declaration.
#pragma once

#include <cstddef>
#include <cstdint>
#include <memory>
#include <string>

#include "axis_hip.hpp"

namespace axis_hip::ingest {

enum class ConventionKind : std::uint8_t { CF, UGRID, GRIB, Projected, NamedGrid, GridRules };  // grid_descriptor.hpp:37-44
enum class CoordinateSystem : std::uint8_t { SphericalDeg, SphericalRad, Cartesian3D };         // grid_descriptor.hpp:52-56

template <class T>
struct View1 {  // field_view<T, 1>
    T *data = nullptr;
    std::size_t n = 0;
};
template <class T>
struct View2 {  // field_view<T, 2>, layout_left
    T *data = nullptr;
    std::size_t n0 = 0, n1 = 0;
};

struct CfParams {  // :63-74
    std::string grid_mapping_name;
    double earth_radius = 0.0, semi_major = 0.0, semi_minor = 0.0, inverse_flattening = 0.0;
};
struct UgridParams {  // :77-84
    int topology_dimension = 2;
    int start_index = 0;
    std::string mesh_name;
};
struct GribParams {  // :87-100
    std::string grid_type;
    std::int64_t ni = 0, nj = 0, gaussian_n = 0;
    std::uint8_t scanning_mode = 0;
    int earth_shape_code = 0;
};
struct ProjectedParams {  // :103-106
    std::string proj_string;
};
struct NamedGridParams {  // :109-112
    std::string name;
};
struct GridRulesParams {  // :116-135
    std::string kind;
    double min_x = -180.0, max_x = 180.0, min_y = -90.0, max_y = 90.0, r_x = 1.0, r_y = 1.0;
    std::int64_t gaussian_n = 0;
    std::string proj_string;
};
struct BufferViews {  // :145-177
    View1<const double> center_x, center_y, corner_x, corner_y;
    View2<const double> node_coords;
    View1<const std::int64_t> conn_offsets, conn_indices;
    View1<const double> cell_areas;
    View1<const int> cell_mask;
    std::size_t ni = 0, nj = 0;
};
struct GridDescriptor {  // :192-216
    ConventionKind kind = ConventionKind::CF;
    CoordinateSystem coord_system = CoordinateSystem::SphericalDeg;
    CfParams cf;
    UgridParams ugrid;
    GribParams grib;
    ProjectedParams projected;
    NamedGridParams named_grid;
    GridRulesParams grid_rules;
    BufferViews buffers;
};

static_assert(sizeof(std::string) == 32, "libstdc++ std::string expected");
static_assert(offsetof(GridDescriptor, cf) == 8 && offsetof(GridDescriptor, ugrid) == 72 && offsetof(GridDescriptor, grib) == 112 &&
                  offsetof(GridDescriptor, projected) == 176 && offsetof(GridDescriptor, named_grid) == 208 &&
                  offsetof(GridDescriptor, grid_rules) == 240 && offsetof(GridDescriptor, buffers) == 360 && sizeof(GridDescriptor) == 528,
              "GridDescriptor layout");
static_assert(offsetof(BufferViews, node_coords) == 64 && offsetof(BufferViews, conn_offsets) == 88 && offsetof(BufferViews, cell_mask) == 136 &&
                  offsetof(BufferViews, ni) == 152,
              "BufferViews layout");

/// MeshFactory::from_descriptor (mesh_factory.cpp:396-431): validates the
/// descriptor with the reference's messages and builds the mesh for the UGRID,
/// CF and GRIB kinds and for GridRules of kind RegularLatLon; the Projected
/// and NamedGrid kinds (PROJ and the named-grid registry) and the other rule
/// kinds are refused with InvalidArgument.
std::shared_ptr<Mesh> mesh_from_descriptor(const GridDescriptor &d);

/// RuleGenerator's RegularLatLon (rule_generator.cpp:105-169): nodes at
/// min + i r, ni = floor((max_x - min_x) / r_x), nj likewise, cell i + j ni with
/// corners BL, BR, TR, TL.
std::shared_ptr<Mesh> mesh_regular_latlon(double min_x, double max_x, double min_y, double max_y, double r_x, double r_y);

/// StructuredGrid(ni, nj, centres).to_unstructured() (structured_grid.cpp:23-272):
/// corners given ((ni + 1)(nj + 1) each) or synthesized from the centres
/// (mean of the up to four surrounding centres, periodic in longitude when
/// ni times the first row's spacing spans 350-370 degrees); node ci + cj (ni + 1),
/// cell i + j ni with corners BL, BR, TR, TL.
std::shared_ptr<Mesh> mesh_from_structured(std::size_t ni, std::size_t nj, const double *center_lon, std::size_t n_center_lon,
                                           const double *center_lat, std::size_t n_center_lat, const double *corner_lon,
                                           std::size_t n_corner_lon, const double *corner_lat, std::size_t n_corner_lat, axis_hip::CoordinateSystem cs);

}  // namespace axis_hip::ingest
