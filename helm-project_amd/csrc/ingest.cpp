// ingest.cpp -- meshes from the reference's ingest contract: GridDescriptor
// (MeshFactory::from_descriptor, mesh_factory.cpp:396-431), the RegularLatLon
// rules (rule_generator.cpp:105-169) and StructuredGrid::to_unstructured
// (structured_grid.cpp:209-272).  Host code: the arrays are built on the host
// once and copied to the device by the Mesh constructor.
#include <cmath>
#include <string>
#include <vector>

#include "grid_descriptor.hpp"

namespace axis_hip::ingest {

namespace {

axis_hip::CoordinateSystem coord_of(CoordinateSystem c) {
    switch (c) {
        case CoordinateSystem::SphericalDeg: return axis_hip::CoordinateSystem::SphericalDeg;
        case CoordinateSystem::SphericalRad: return axis_hip::CoordinateSystem::SphericalRad;
        case CoordinateSystem::Cartesian3D: return axis_hip::CoordinateSystem::Cartesian3D;
    }
    throw InvalidArgument("MeshFactory::from_descriptor: unknown CoordinateSystem value (" + std::to_string(static_cast<int>(c)) + ")");
}

bool empty(const View1<const double> &v) { return v.data == nullptr || v.n == 0; }

// centre-array checks shared by the CF, GRIB and Projected kinds (mesh_factory.cpp:56-92, 117-157)
void check_centres(const char *kind, const BufferViews &b, std::size_t expected) {
    const std::string pre = std::string("MeshFactory::from_descriptor(") + kind + "): ";
    if (empty(b.center_x)) throw InvalidArgument(pre + "null/empty required buffer 'center_x'");
    if (empty(b.center_y)) throw InvalidArgument(pre + "null/empty required buffer 'center_y'");
    if (b.center_x.n != expected)
        throw InvalidArgument(pre + "inconsistent buffer extents — center_x.extent(0)=" + std::to_string(b.center_x.n) +
                              " but ni*nj=" + std::to_string(expected));
    if (b.center_y.n != expected)
        throw InvalidArgument(pre + "inconsistent buffer extents — center_y.extent(0)=" + std::to_string(b.center_y.n) +
                              " but ni*nj=" + std::to_string(expected));
}

std::shared_ptr<Mesh> structured_of(std::size_t ni, std::size_t nj, const BufferViews &b, CoordinateSystem cs) {
    const bool corners = !empty(b.corner_x) && !empty(b.corner_y);
    return mesh_from_structured(ni, nj, b.center_x.data, b.center_x.n, b.center_y.data, b.center_y.n, corners ? b.corner_x.data : nullptr,
                                corners ? b.corner_x.n : 0, corners ? b.corner_y.data : nullptr, corners ? b.corner_y.n : 0, coord_of(cs));
}

}  // namespace

std::shared_ptr<Mesh> mesh_regular_latlon(double min_x, double max_x, double min_y, double max_y, double r_x, double r_y) {
    // validate_bbox / validate_resolution (rule_generator.cpp:29-47)
    if (!(max_x > min_x))
        throw InvalidArgument("RuleGenerator: max_x (" + std::to_string(max_x) + ") must be greater than min_x (" + std::to_string(min_x) + ")");
    if (!(max_y > min_y))
        throw InvalidArgument("RuleGenerator: max_y (" + std::to_string(max_y) + ") must be greater than min_y (" + std::to_string(min_y) + ")");
    if (!(r_x > 0.0)) throw InvalidArgument("RuleGenerator: r_x must be positive, got " + std::to_string(r_x));
    if (!(r_y > 0.0)) throw InvalidArgument("RuleGenerator: r_y must be positive, got " + std::to_string(r_y));
    const auto ni = static_cast<index_t>(std::floor((max_x - min_x) / r_x));
    const auto nj = static_cast<index_t>(std::floor((max_y - min_y) / r_y));
    if (ni == 0 || nj == 0)
        throw InvalidArgument("RuleGenerator: RegularLatLon resolution produces zero cells (ni=" + std::to_string(ni) + ", nj=" + std::to_string(nj) +
                              ")");
    const index_t nn = (ni + 1) * (nj + 1), nc = ni * nj, st = ni + 1;
    std::vector<double> coords(2 * nn);
    for (index_t j = 0; j <= nj; ++j)
        for (index_t i = 0; i <= ni; ++i) {
            coords[j * st + i] = min_x + static_cast<double>(i) * r_x;
            coords[nn + j * st + i] = min_y + static_cast<double>(j) * r_y;
        }
    std::vector<index_t> off(nc + 1), ind(4 * nc);
    for (index_t c = 0; c <= nc; ++c) off[c] = 4 * c;
    for (index_t j = 0; j < nj; ++j)
        for (index_t i = 0; i < ni; ++i) {
            index_t *q = &ind[4 * (i + j * ni)];
            q[0] = j * st + i;
            q[1] = j * st + i + 1;
            q[2] = (j + 1) * st + i + 1;
            q[3] = (j + 1) * st + i;
        }
    return std::make_shared<Mesh>(coords.data(), nn, 2, off.data(), nc, ind.data(), axis_hip::CoordinateSystem::SphericalDeg, nullptr, nullptr);
}

std::shared_ptr<Mesh> mesh_from_structured(std::size_t ni, std::size_t nj, const double *center_lon, std::size_t n_center_lon,
                                           const double *center_lat, std::size_t n_center_lat, const double *corner_lon,
                                           std::size_t n_corner_lon, const double *corner_lat, std::size_t n_corner_lat,
                                           axis_hip::CoordinateSystem cs) {
    // StructuredGrid constructor and set_corners checks (structured_grid.cpp:23-83)
    if (ni == 0 || nj == 0) throw InvalidArgument("StructuredGrid: ni and nj must be positive");
    const std::size_t nc = ni * nj, nip1 = ni + 1, nn = (ni + 1) * (nj + 1);
    if (n_center_lon != nc)
        throw InvalidArgument("StructuredGrid: center_lon extent (" + std::to_string(n_center_lon) + ") does not match ni*nj (" + std::to_string(nc) +
                              ")");
    if (n_center_lat != nc)
        throw InvalidArgument("StructuredGrid: center_lat extent (" + std::to_string(n_center_lat) + ") does not match ni*nj (" + std::to_string(nc) +
                              ")");
    std::vector<double> coords(2 * nn);
    double *nlon = coords.data(), *nlat = coords.data() + nn;
    if (corner_lon != nullptr) {
        if (n_corner_lon != nn)
            throw InvalidArgument("StructuredGrid::set_corners: corner_lon extent (" + std::to_string(n_corner_lon) +
                                  ") does not match (ni+1)*(nj+1) (" + std::to_string(nn) + ")");
        if (n_corner_lat != nn)
            throw InvalidArgument("StructuredGrid::set_corners: corner_lat extent (" + std::to_string(n_corner_lat) +
                                  ") does not match (ni+1)*(nj+1) (" + std::to_string(nn) + ")");
        for (std::size_t q = 0; q < nn; ++q) {
            nlon[q] = corner_lon[q];
            nlat[q] = corner_lat[q];
        }
    } else {
        // corner (ci, cj) = mean of the centres of cells (ci-1 .. ci, cj-1 .. cj)
        // that exist, visited row below first, left cell first; in a periodic
        // row the cell across the seam is taken with its longitude moved by
        // -360 (left) or +360 (right)
        const bool periodic = ni > 1 && [&] {
            const double span = static_cast<double>(ni) * std::abs(center_lon[1] - center_lon[0]);
            return span > 350.0 && span < 370.0;
        }();
        for (std::size_t cj = 0; cj <= nj; ++cj)
            for (std::size_t ci = 0; ci <= ni; ++ci) {
                double sx = 0.0, sy = 0.0;
                int cnt = 0;
                for (int dj = -1; dj <= 0; ++dj) {
                    const long long j = static_cast<long long>(cj) + dj;
                    for (int di = -1; di <= 0; ++di) {
                        long long i = static_cast<long long>(ci) + di;
                        double shift = 0.0;
                        bool ok;
                        if (periodic) {
                            if (i < 0) {
                                i = static_cast<long long>(ni) - 1;
                                shift = -360.0;
                            } else if (i >= static_cast<long long>(ni)) {
                                i = 0;
                                shift = 360.0;
                            }
                            ok = true;
                        } else {
                            ok = i >= 0 && i < static_cast<long long>(ni);
                        }
                        if (!ok || j < 0 || j >= static_cast<long long>(nj)) continue;
                        const std::size_t c = static_cast<std::size_t>(i) + static_cast<std::size_t>(j) * ni;
                        sx += center_lon[c] + shift;
                        sy += center_lat[c];
                        ++cnt;
                    }
                }
                nlon[ci + cj * nip1] = sx / static_cast<double>(cnt);
                nlat[ci + cj * nip1] = sy / static_cast<double>(cnt);
            }
    }
    std::vector<index_t> off(nc + 1), ind(4 * nc);
    for (std::size_t c = 0; c <= nc; ++c) off[c] = static_cast<index_t>(4 * c);
    for (std::size_t j = 0; j < nj; ++j)
        for (std::size_t i = 0; i < ni; ++i) {
            index_t *q = &ind[4 * (i + j * ni)];
            q[0] = static_cast<index_t>(i + j * nip1);
            q[1] = static_cast<index_t>(i + 1 + j * nip1);
            q[2] = static_cast<index_t>(i + 1 + (j + 1) * nip1);
            q[3] = static_cast<index_t>(i + (j + 1) * nip1);
        }
    return std::make_shared<Mesh>(coords.data(), static_cast<index_t>(nn), 2, off.data(), static_cast<index_t>(nc), ind.data(), cs, nullptr, nullptr);
}

std::shared_ptr<Mesh> mesh_from_descriptor(const GridDescriptor &d) {
    const BufferViews &b = d.buffers;
    switch (d.kind) {
        case ConventionKind::CF: {  // validate_cf, build_from_cf (mesh_factory.cpp:57-92, 263-285)
            if (b.ni == 0) throw InvalidArgument("MeshFactory::from_descriptor(CF): missing required field 'ni' (must be > 0)");
            if (b.nj == 0) throw InvalidArgument("MeshFactory::from_descriptor(CF): missing required field 'nj' (must be > 0)");
            check_centres("CF", b, b.ni * b.nj);
            return structured_of(b.ni, b.nj, b, d.coord_system);
        }
        case ConventionKind::UGRID: {  // validate_ugrid, build_from_ugrid (mesh_factory.cpp:95-115, 287-312)
            if (b.node_coords.data == nullptr || b.node_coords.n0 == 0)
                throw InvalidArgument("MeshFactory::from_descriptor(UGRID): null/empty required buffer 'node_coords'");
            if (b.conn_offsets.data == nullptr || b.conn_offsets.n == 0)
                throw InvalidArgument("MeshFactory::from_descriptor(UGRID): null/empty required buffer 'conn_offsets'");
            if (b.conn_indices.data == nullptr || b.conn_indices.n == 0)
                throw InvalidArgument("MeshFactory::from_descriptor(UGRID): null/empty required buffer 'conn_indices'");
            if (b.node_coords.n1 < 2)
                throw InvalidArgument("MeshFactory::from_descriptor(UGRID): inconsistent buffer extents — node_coords.extent(1)=" +
                                      std::to_string(b.node_coords.n1) + " but ndim must be >= 2");
            const index_t n_cells = static_cast<index_t>(b.conn_offsets.n) - 1;
            if (n_cells >= 0 && b.conn_offsets.data[n_cells] > static_cast<index_t>(b.conn_indices.n))
                throw InvalidArgument("MeshFactory::from_descriptor(UGRID): conn_offsets reach past conn_indices");
            const bool areas = b.cell_areas.data != nullptr && b.cell_areas.n > 0;
            const bool mask = b.cell_mask.data != nullptr && b.cell_mask.n > 0;
            if (areas && static_cast<index_t>(b.cell_areas.n) < n_cells) throw InvalidArgument("MeshFactory::from_descriptor(UGRID): cell_areas shorter than n_cells");
            if (mask && static_cast<index_t>(b.cell_mask.n) < n_cells) throw InvalidArgument("MeshFactory::from_descriptor(UGRID): cell_mask shorter than n_cells");
            return std::make_shared<Mesh>(b.node_coords.data, static_cast<index_t>(b.node_coords.n0), static_cast<index_t>(b.node_coords.n1),
                                          b.conn_offsets.data, n_cells, b.conn_indices.data, coord_of(d.coord_system), areas ? b.cell_areas.data : nullptr,
                                          mask ? b.cell_mask.data : nullptr);
        }
        case ConventionKind::GRIB: {  // validate_grib, build_from_grib (mesh_factory.cpp:118-157, 314-339)
            const GribParams &g = d.grib;
            if (g.grid_type.empty()) throw InvalidArgument("MeshFactory::from_descriptor(GRIB): missing required field 'grid_type'");
            if (g.ni <= 0) throw InvalidArgument("MeshFactory::from_descriptor(GRIB): missing required field 'ni' (must be > 0)");
            if (g.nj <= 0) throw InvalidArgument("MeshFactory::from_descriptor(GRIB): missing required field 'nj' (must be > 0)");
            const std::size_t ni = static_cast<std::size_t>(g.ni), nj = static_cast<std::size_t>(g.nj);
            check_centres("GRIB", b, ni * nj);
            return structured_of(ni, nj, b, d.coord_system);
        }
        case ConventionKind::Projected:
            if (d.projected.proj_string.empty())
                throw InvalidArgument("MeshFactory::from_descriptor(Projected): missing required field 'proj_string' (must be non-empty)");
            throw InvalidArgument("MeshFactory::from_descriptor(Projected): projected grids need PROJ, which this library does not link");
        case ConventionKind::NamedGrid:
            if (d.named_grid.name.empty())
                throw InvalidArgument("MeshFactory::from_descriptor(NamedGrid): missing required field 'name' (must be non-empty)");
            throw InvalidArgument("MeshFactory::from_descriptor(NamedGrid): named grids are not built ('" + d.named_grid.name + "')");
        case ConventionKind::GridRules: {  // validate_grid_rules, RuleGenerator::generate (rule_generator.cpp:470-490)
            const GridRulesParams &r = d.grid_rules;
            if (r.kind.empty()) throw InvalidArgument("MeshFactory::from_descriptor(GridRules): missing required field 'kind' (must be non-empty)");
            if (r.kind == "RegularLatLon") return mesh_regular_latlon(r.min_x, r.max_x, r.min_y, r.max_y, r.r_x, r.r_y);
            throw InvalidArgument("RuleGenerator: rule kind '" + r.kind + "' is not built (RegularLatLon only)");
        }
    }
    throw InvalidArgument("MeshFactory::from_descriptor: unknown ConventionKind value (" + std::to_string(static_cast<int>(d.kind)) + ")");
}

}  // namespace axis_hip::ingest
