// axis_oracle.cpp -- CPU restatement of the AXIS regrid host path.
//
// TEST INFRASTRUCTURE ONLY.  This file is the checker for the HIP product
// library: only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
// leg load it.  The product never links, loads or calls it.
//
// It restates, in plain C++20 (g++ -O3 -fopenmp, no fast-math, no FMA
// contraction), the reference host path of libs/axis (paths relative to the
// reference repository root):
//   * SphericalPolygon::area / SphericalClipper      spherical_clipper.hpp:29-336
//   * lonlat_to_xyz                                  spherical_geometry.hpp:80-83
//   * extract_cell_polygon_spherical(_cached)        weight_generator.cpp:441-520
//   * get_cell_areas_spherical                       weight_generator.cpp:523-552
//   * compute_cell_aabbs_3d                          weight_generator.cpp:198-262
//   * cap filter                                     spherical_cap_filter.hpp:63-251
//   * NodeTrigCache                                  trig_cache.hpp:179-289
//   * detect_regular_grid / detect_rectilinear_grid  regular_grid_detector.hpp:64-299
//   * DegenerateCellHandler::classify / scan         degenerate_cell_handler.hpp:115-367
//   * morton_sort_indices                            morton_sort.hpp:43-176
//   * generate() checks + dispatch                   weight_generator.cpp:1887-2019
//   * generate_conservative(_impl) (great circle)    weight_generator.cpp:3246-3770
//   * generate_bilinear_rect_nonuniform              weight_generator_bilinear_rect_nonuniform.cpp:22-149
//   * InterpolationMatrix::to_csr                    interpolation_matrix.hpp:180-255
//   * apply / batch_apply (CSR, sequential rows)     apply.hpp:61-343
//   * nan_batch_apply, masked apply                  apply.hpp:359-484
//   * check_conservation                             conservation.cpp:23-103
//
// Broad phase: the reference queries an ArborX v2.1 BVH with float boxes
// (third party, not in the reference tree).  The oracle finds exactly the
// same SET of intersecting float-box pairs, either by brute force or with its
// own median-split BVH; candidate order is irrelevant after to_csr except for
// the order of the frac_b sums, which the oracle takes in ascending source
// index.
//
// Every function is exported with a C ABI (prefix `om_`) for ctypes.

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <numeric>
#include <stdexcept>
#include <string>
#include <vector>
#include <omp.h>

namespace {

using i64 = std::int64_t;

constexpr double kPi = 3.14159265358979323846;
constexpr double kDeg2Rad = kPi / 180.0;  // spherical_geometry.hpp:36, spherical_cap_filter.hpp:40, trig_cache.hpp:35

struct V3 {
    double x, y, z;
};
inline double dot(const V3 &a, const V3 &b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline V3 cross(const V3 &a, const V3 &b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
inline double lsq(const V3 &v) { return v.x * v.x + v.y * v.y + v.z * v.z; }
inline V3 nrm(const V3 &v) {  // spherical_clipper.hpp:60-66
    double len = std::sqrt(lsq(v));
    if (len < 1.0e-300) return {0.0, 0.0, 0.0};
    double inv = 1.0 / len;
    return {v.x * inv, v.y * inv, v.z * inv};
}
inline V3 lonlat_xyz(double lon_rad, double lat_rad) {  // spherical_geometry.hpp:80-83
    double cl = std::cos(lat_rad);
    return {cl * std::cos(lon_rad), cl * std::sin(lon_rad), std::sin(lat_rad)};
}

// Fixed capacity polygon, spherical_clipper.hpp:100-151.
struct Poly {
    V3 v[32];
    int n = 0;
    void push(const V3 &p) {
        if (n < 32) v[n++] = p;
    }
    // Van Oosterom-Strackee fan from vertex 0, spherical_clipper.hpp:112-137.
    double area() const {
        if (n < 3) return 0.0;
        double total = 0.0;
        const V3 &a = v[0];
        for (int i = 1; i < n - 1; ++i) {
            const V3 &b = v[i];
            const V3 &c = v[i + 1];
            double ab = dot(a, b), ac = dot(a, c), bc = dot(b, c);
            double num = dot(a, cross(b, c));
            double den = 1.0 + ab + ac + bc;
            total += 2.0 * std::atan2(num, den);
        }
        return std::fabs(total);
    }
};

// spherical_clipper.hpp:263-286
V3 gc_intersection(const V3 &v1, const V3 &v2, const V3 &ea, const V3 &eb) {
    V3 n1 = cross(ea, eb);
    V3 n2 = cross(v1, v2);
    V3 d = cross(n1, n2);
    if (lsq(d) < 1.0e-30) {
        V3 mid{v1.x + v2.x, v1.y + v2.y, v1.z + v2.z};
        return nrm(mid);
    }
    V3 cand = nrm(d);
    V3 mid_dir = nrm(V3{v1.x + v2.x, v1.y + v2.y, v1.z + v2.z});
    if (dot(cand, mid_dir) < 0.0) cand = {-cand.x, -cand.y, -cand.z};
    return cand;
}

// Sutherland-Hodgman of subject against every clip edge,
// spherical_clipper.hpp:174-222 and :301-335.
Poly clip(const Poly &subject, const Poly &cp) {
    if (subject.n < 3 || cp.n < 3) return Poly{};
    Poly a = subject, b;
    for (int ci = 0; ci < cp.n; ++ci) {
        const V3 &ea = cp.v[ci];
        const V3 &eb = cp.v[(ci + 1) % cp.n];
        V3 normal = cross(ea, eb);
        if (lsq(normal) < 1.0e-30) continue;
        b.n = 0;
        if (a.n >= 1) {
            constexpr double eps = -1.0e-14;
            for (int i = 0; i < a.n; ++i) {
                int p = (i + a.n - 1) % a.n;
                const V3 &cur = a.v[i];
                const V3 &prv = a.v[p];
                double dc = dot(normal, cur), dp = dot(normal, prv);
                bool ci_in = dc >= eps, pi_in = dp >= eps;
                if (ci_in) {
                    if (!pi_in) b.push(gc_intersection(prv, cur, ea, eb));
                    b.push(cur);
                } else if (pi_in) {
                    b.push(gc_intersection(prv, cur, ea, eb));
                }
            }
        }
        a.n = b.n;
        for (int i = 0; i < b.n; ++i) a.v[i] = b.v[i];
        if (a.n < 3) return Poly{};
    }
    return a;
}

// ---------------------------------------------------------------- meshes
struct Mesh {
    const double *coords;  // layout_left [n_nodes, 2]: lon block then lat block
    i64 n_nodes;
    const i64 *offsets;  // [n_cells + 1]
    const i64 *indices;
    i64 n_cells;
    int coord_sys;          // 0 SphericalDeg, 1 SphericalRad, 2 Cartesian3D
    const double *areas;    // optional [n_cells] or null
    const int32_t *mask;    // optional [n_cells] or null
    double lon(i64 n) const { return coords[n]; }
    double lat(i64 n) const { return coords[n + n_nodes]; }
};

// extract_cell_polygon_spherical, weight_generator.cpp:441-480 (degree and
// radian branches; the Cartesian branch is outside the great-circle path).
Poly cell_poly(const Mesh &m, i64 c) {
    Poly p;
    for (i64 k = m.offsets[c]; k < m.offsets[c + 1]; ++k) {
        i64 nd = m.indices[k];
        double c0 = m.lon(nd), c1 = m.lat(nd);
        if (m.coord_sys == 0)
            p.push(lonlat_xyz(c0 * kDeg2Rad, c1 * kDeg2Rad));
        else
            p.push(lonlat_xyz(c0, c1));
    }
    return p;
}

// ------------------------------------------------ regular grid detection
struct RegularInfo {
    bool is_regular = false;
    double lon_min = 0, lon_max = 0, delta_lon = 0, lat_min = 0, lat_max = 0, delta_lat = 0;
    i64 ni = 0, nj = 0;
};
struct RectInfo {
    bool is_rect = false;
    i64 ni = 0, nj = 0;
    std::vector<double> ulon, ulat;
};

bool all_quads(const Mesh &m) {
    for (i64 c = 0; c < m.n_cells; ++c)
        if (m.offsets[c + 1] - m.offsets[c] != 4) return false;
    return true;
}
// regular_grid_detector.hpp:100-112: tolerance unique over sorted values,
// compared with the last kept value.
std::vector<double> unique_tol(std::vector<double> s) {
    std::sort(s.begin(), s.end());
    std::vector<double> u;
    if (s.empty()) return u;
    u.push_back(s[0]);
    for (size_t i = 1; i < s.size(); ++i)
        if (std::abs(s[i] - u.back()) > 1.0e-12) u.push_back(s[i]);
    return u;
}
void unique_axes(const Mesh &m, std::vector<double> &ul, std::vector<double> &ua) {
    std::vector<double> lons(m.coords, m.coords + m.n_nodes), lats(m.coords + m.n_nodes, m.coords + 2 * m.n_nodes);
    ul = unique_tol(std::move(lons));
    ua = unique_tol(std::move(lats));
}
// regular_grid_detector.hpp:165-299
RegularInfo detect_regular(const Mesh &m) {
    RegularInfo info;
    if (m.n_cells == 0 || !all_quads(m)) return info;
    std::vector<double> ul, ua;
    unique_axes(m, ul, ua);
    if (ul.size() < 2 || ua.size() < 2) return info;
    i64 ni = (i64)ul.size() - 1, nj = (i64)ua.size() - 1;
    if (ni * nj != m.n_cells) return info;
    auto deltas = [](const std::vector<double> &u) {
        std::vector<double> d(u.size() - 1);
        for (size_t i = 0; i + 1 < u.size(); ++i) d[i] = u[i + 1] - u[i];
        return d;
    };
    auto uniform = [](const std::vector<double> &d) {
        if (d.empty()) return false;
        double mn = d[0], mx = d[0], sum = 0.0;
        for (double x : d) {
            mn = std::min(mn, x);
            mx = std::max(mx, x);
            sum += x;
        }
        if (mn <= 0.0) return false;
        double mean = sum / (double)d.size();
        return (mx - mn) < 1.0e-10 * mean;
    };
    auto dl = deltas(ul), da = deltas(ua);
    if (!uniform(dl) || !uniform(da)) return info;
    double sl = 0.0, sa = 0.0;
    for (double x : dl) sl += x;
    for (double x : da) sa += x;
    info.is_regular = true;
    info.lon_min = ul.front();
    info.lon_max = ul.back();
    info.delta_lon = sl / (double)ni;
    info.lat_min = ua.front();
    info.lat_max = ua.back();
    info.delta_lat = sa / (double)nj;
    info.ni = ni;
    info.nj = nj;
    return info;
}
// regular_grid_detector.hpp:64-144
RectInfo detect_rect(const Mesh &m) {
    RectInfo info;
    if (m.n_cells == 0 || !all_quads(m)) return info;
    std::vector<double> ul, ua;
    unique_axes(m, ul, ua);
    if (ul.size() < 2 || ua.size() < 2) return info;
    i64 ni = (i64)ul.size() - 1, nj = (i64)ua.size() - 1;
    if (ni * nj != m.n_cells) return info;
    info.is_rect = true;
    info.ni = ni;
    info.nj = nj;
    info.ulon = std::move(ul);
    info.ulat = std::move(ua);
    return info;
}

// --------------------------------------------------------- trig cache
// trig_cache.hpp:210-289: node tables at (min + i*delta) * deg2rad.
struct NodeTrig {
    bool valid = false;
    i64 ni = 0, nj = 0;
    std::vector<double> slon, clon, slat, clat;
};
NodeTrig build_node_trig(const RegularInfo &g) {
    NodeTrig t;
    if (!g.is_regular || g.ni == 0 || g.nj == 0) return t;
    t.ni = g.ni;
    t.nj = g.nj;
    t.slon.resize(g.ni + 1);
    t.clon.resize(g.ni + 1);
    t.slat.resize(g.nj + 1);
    t.clat.resize(g.nj + 1);
    for (i64 i = 0; i <= g.ni; ++i) {
        double r = (g.lon_min + (double)i * g.delta_lon) * kDeg2Rad;
        t.slon[i] = std::sin(r);
        t.clon[i] = std::cos(r);
    }
    for (i64 j = 0; j <= g.nj; ++j) {
        double r = (g.lat_min + (double)j * g.delta_lat) * kDeg2Rad;
        t.slat[j] = std::sin(r);
        t.clat[j] = std::cos(r);
    }
    t.valid = true;
    return t;
}
// extract_cell_polygon_spherical_cached, weight_generator.cpp:483-520 (BL, BR, TR, TL)
Poly cell_poly_cached(const NodeTrig &t, i64 c) {
    i64 ci = c % t.ni, cj = c / t.ni;
    auto node = [&](i64 a, i64 b) { return V3{t.clat[b] * t.clon[a], t.clat[b] * t.slon[a], t.slat[b]}; };
    Poly p;
    p.push(node(ci, cj));
    p.push(node(ci + 1, cj));
    p.push(node(ci + 1, cj + 1));
    p.push(node(ci, cj + 1));
    return p;
}

// ------------------------------------------------------ degenerate scan
// degenerate_cell_handler.hpp:115-142 and helpers :248-367.
int count_unique(const Poly &p, double thr) {
    if (p.n < 1) return 0;
    double t2 = thr * thr;
    int u = 1;
    for (int i = 1; i < p.n; ++i) {
        bool dup = false;
        for (int j = 0; j < i; ++j) {
            V3 d{p.v[i].x - p.v[j].x, p.v[i].y - p.v[j].y, p.v[i].z - p.v[j].z};
            if (lsq(d) < t2) {
                dup = true;
                break;
            }
        }
        if (!dup) ++u;
    }
    return u;
}
bool edges_cross(const V3 &a, const V3 &b, const V3 &c, const V3 &d) {
    V3 nab = cross(a, b);
    double dc = dot(nab, c), dd = dot(nab, d);
    if (dc * dd >= 0.0) return false;
    V3 ncd = cross(c, d);
    double da = dot(ncd, a), db = dot(ncd, b);
    if (da * db >= 0.0) return false;
    return true;
}
bool self_intersecting(const Poly &p) {
    if (p.n < 4) return false;
    for (int i = 0; i < p.n; ++i) {
        int in = (i + 1) % p.n;
        for (int j = i + 2; j < p.n; ++j) {
            int jn = (j + 1) % p.n;
            if (jn == i) continue;
            if (edges_cross(p.v[i], p.v[in], p.v[j], p.v[jn])) return true;
        }
    }
    return false;
}
int classify(const Poly &p) {
    if (p.n < 3) return 1;
    if (p.area() < 1e-25) return 1;
    if (count_unique(p, 1e-15) < 3) return 2;
    if (self_intersecting(p)) return 3;
    return 0;
}
// scan() converts (lon, lat) as DEGREES whatever the coordinate system,
// degenerate_cell_handler.hpp:198-215.
int classify_cell(const Mesh &m, i64 c) {
    Poly p;
    for (i64 k = m.offsets[c]; k < m.offsets[c + 1]; ++k) {
        i64 nd = m.indices[k];
        double lon = m.lon(nd) * kDeg2Rad, lat = m.lat(nd) * kDeg2Rad;
        double cl = std::cos(lat);
        p.push(V3{cl * std::cos(lon), cl * std::sin(lon), std::sin(lat)});
    }
    return classify(p);
}

// ----------------------------------------------------------- cap data
// spherical_cap_filter.hpp:85-251
void cap_data(const Mesh &m, std::vector<V3> &cen, std::vector<double> &rad) {
    cen.resize(m.n_cells);
    rad.resize(m.n_cells);
    const bool deg = (m.coord_sys == 0);
#pragma omp parallel for schedule(static)
    for (i64 c = 0; c < m.n_cells; ++c) {
        i64 b = m.offsets[c], e = m.offsets[c + 1];
        int nv = (int)(e - b);
        V3 ce{0, 0, 0};
        if (nv >= 1) {
            double sx = 0, sy = 0, sz = 0;
            for (int v = 0; v < nv; ++v) {
                i64 nd = m.indices[b + v];
                double lo = m.lon(nd), la = m.lat(nd);
                if (deg) {
                    lo *= kDeg2Rad;
                    la *= kDeg2Rad;
                }
                V3 p = lonlat_xyz(lo, la);
                sx += p.x;
                sy += p.y;
                sz += p.z;
            }
            double inv = 1.0 / (double)nv;
            ce = nrm(V3{sx * inv, sy * inv, sz * inv});
        }
        double mx = 0.0;
        for (int v = 0; v < nv; ++v) {
            i64 nd = m.indices[b + v];
            double lo = m.lon(nd), la = m.lat(nd);
            if (deg) {
                lo *= kDeg2Rad;
                la *= kDeg2Rad;
            }
            V3 p = lonlat_xyz(lo, la);
            double d = dot(ce, p);
            d = std::fmin(1.0, std::fmax(-1.0, d));
            double ang = std::acos(d);
            if (ang > mx) mx = ang;
        }
        cen[c] = ce;
        rad[c] = mx;
    }
}
inline bool cap_rejects(const V3 &cs, double rs, const V3 &cd, double rd) {  // :190-197
    double cd_ = dot(cs, cd);
    cd_ = std::fmin(1.0, std::fmax(-1.0, cd_));
    return std::acos(cd_) > (rs + rd);
}

// ------------------------------------------------------ float AABBs
// compute_cell_aabbs_3d, weight_generator.cpp:198-262
struct FBox {
    float lo[3], hi[3];
};
std::vector<FBox> aabbs3d(const Mesh &m) {
    std::vector<FBox> boxes(m.n_cells);
#pragma omp parallel for schedule(static)
    for (i64 c = 0; c < m.n_cells; ++c) {
        double mn[3] = {std::numeric_limits<double>::max(), std::numeric_limits<double>::max(), std::numeric_limits<double>::max()};
        double mx[3] = {-std::numeric_limits<double>::max(), -std::numeric_limits<double>::max(), -std::numeric_limits<double>::max()};
        for (i64 k = m.offsets[c]; k < m.offsets[c + 1]; ++k) {
            i64 nd = m.indices[k];
            double lon = m.lon(nd), lat = m.lat(nd);
            if (m.coord_sys == 0) {
                lon = lon * kPi / 180.0;
                lat = lat * kPi / 180.0;
            }
            double p[3] = {std::cos(lat) * std::cos(lon), std::cos(lat) * std::sin(lon), std::sin(lat)};
            for (int a = 0; a < 3; ++a) {
                mn[a] = std::min(mn[a], p[a]);
                mx[a] = std::max(mx[a], p[a]);
            }
        }
        double ext = std::max({mx[0] - mn[0], mx[1] - mn[1], mx[2] - mn[2]});
        double eps = std::min(std::max(0.25 * ext, 1.0e-6), 0.02);
        for (int a = 0; a < 3; ++a) {
            boxes[c].lo[a] = (float)(mn[a] - eps);
            boxes[c].hi[a] = (float)(mx[a] + eps);
        }
    }
    return boxes;
}
inline bool box_hit(const FBox &a, const FBox &b) {
    for (int k = 0; k < 3; ++k)
        if (a.hi[k] < b.lo[k] || b.hi[k] < a.lo[k]) return false;
    return true;
}

// Median-split BVH over float boxes (the oracle's own broad phase; finds
// every box pair that intersects, same set as the reference's ArborX query).
struct Bvh {
    struct Node {
        FBox box;
        int left, right;  // children, or -1
        i64 begin, end;   // leaf range in perm
    };
    std::vector<Node> nodes;
    std::vector<i64> perm;
    const std::vector<FBox> *boxes;
    int build(i64 b, i64 e, std::vector<float> &cx, int depth) {
        Node nd;
        nd.left = nd.right = -1;
        nd.begin = b;
        nd.end = e;
        FBox bb;
        for (int a = 0; a < 3; ++a) {
            bb.lo[a] = std::numeric_limits<float>::max();
            bb.hi[a] = -std::numeric_limits<float>::max();
        }
        for (i64 i = b; i < e; ++i) {
            const FBox &x = (*boxes)[perm[i]];
            for (int a = 0; a < 3; ++a) {
                bb.lo[a] = std::min(bb.lo[a], x.lo[a]);
                bb.hi[a] = std::max(bb.hi[a], x.hi[a]);
            }
        }
        nd.box = bb;
        int id = (int)nodes.size();
        nodes.push_back(nd);
        if (e - b > 8 && depth < 60) {
            int ax = 0;
            float ext = bb.hi[0] - bb.lo[0];
            for (int a = 1; a < 3; ++a)
                if (bb.hi[a] - bb.lo[a] > ext) {
                    ext = bb.hi[a] - bb.lo[a];
                    ax = a;
                }
            i64 mid = (b + e) / 2;
            std::nth_element(perm.begin() + b, perm.begin() + mid, perm.begin() + e, [&](i64 p, i64 q) {
                const FBox &P = (*boxes)[p], &Q = (*boxes)[q];
                float cp = P.lo[ax] + P.hi[ax], cq = Q.lo[ax] + Q.hi[ax];
                return cp < cq || (cp == cq && p < q);
            });
            int l = build(b, mid, cx, depth + 1);
            int r = build(mid, e, cx, depth + 1);
            nodes[id].left = l;
            nodes[id].right = r;
        }
        return id;
    }
    void init(const std::vector<FBox> &bx) {
        boxes = &bx;
        perm.resize(bx.size());
        std::iota(perm.begin(), perm.end(), 0);
        nodes.reserve(bx.size() / 4 + 16);
        std::vector<float> cx;
        if (!bx.empty()) build(0, (i64)bx.size(), cx, 0);
    }
    void query(const FBox &q, std::vector<i64> &out) const {
        out.clear();
        if (nodes.empty()) return;
        int stack[128];
        int sp = 0;
        stack[sp++] = 0;
        while (sp) {
            const Node &n = nodes[stack[--sp]];
            if (!box_hit(n.box, q)) continue;
            if (n.left < 0) {
                for (i64 i = n.begin; i < n.end; ++i)
                    if (box_hit((*boxes)[perm[i]], q)) out.push_back(perm[i]);
            } else {
                stack[sp++] = n.left;
                stack[sp++] = n.right;
            }
        }
        std::sort(out.begin(), out.end());
    }
};

// ------------------------------------------------------------- Morton
// morton_sort.hpp:43-176 with the centroid rule of
// compute_cell_centroids_xy, weight_generator.cpp:78-138 (spherical branch).
std::vector<i64> morton_order(const Mesh &m) {
    const i64 n = m.n_cells;
    std::vector<double> cx(n), cy(n);
    for (i64 c = 0; c < n; ++c) {
        i64 b = m.offsets[c], e = m.offsets[c + 1];
        int nv = (int)(e - b);
        double sx = 0, sy = 0, sz = 0;
        for (i64 k = b; k < e; ++k) {
            double lon = m.lon(m.indices[k]), lat = m.lat(m.indices[k]);
            if (m.coord_sys == 0) {
                lon = lon * kPi / 180.0;
                lat = lat * kPi / 180.0;
            }
            sx += std::cos(lat) * std::cos(lon);
            sy += std::cos(lat) * std::sin(lon);
            sz += std::sin(lat);
        }
        double inv = nv > 0 ? 1.0 / (double)nv : 0.0;
        double la = std::asin(sz * inv), lo = std::atan2(sy * inv, sx * inv);
        if (lo < 0.0) lo += 2.0 * kPi;
        if (m.coord_sys == 0) {
            lo = lo * 180.0 / kPi;
            la = la * 180.0 / kPi;
        }
        cx[c] = lo;
        cy[c] = la;
    }
    std::vector<i64> perm(n);
    std::iota(perm.begin(), perm.end(), 0);
    if (n == 0) return perm;
    double x0 = 0, x1 = 0, y0 = 0, y1 = 0;  // Kokkos Min/Max reducers start from the identity,
    x0 = *std::min_element(cx.begin(), cx.end());
    x1 = *std::max_element(cx.begin(), cx.end());
    y0 = *std::min_element(cy.begin(), cy.end());
    y1 = *std::max_element(cy.begin(), cy.end());
    auto q = [](double v, double lo, double hi) -> uint32_t {
        const uint32_t N = (1u << 16) - 1;
        if (hi <= lo) return 0;
        double t = (v - lo) / (hi - lo);
        t = std::fmax(0.0, std::fmin(1.0, t));
        return (uint32_t)(t * N);
    };
    auto spread = [](uint32_t v) -> uint64_t {
        uint64_t r = v;
        r = (r | (r << 16)) & 0x0000FFFF0000FFFFull;
        r = (r | (r << 8)) & 0x00FF00FF00FF00FFull;
        r = (r | (r << 4)) & 0x0F0F0F0F0F0F0F0Full;
        r = (r | (r << 2)) & 0x3333333333333333ull;
        r = (r | (r << 1)) & 0x5555555555555555ull;
        return r;
    };
    std::vector<uint64_t> key(n);
    for (i64 c = 0; c < n; ++c) key[c] = spread(q(cx[c], x0, x1)) | (spread(q(cy[c], y0, y1)) << 1);
    std::stable_sort(perm.begin(), perm.end(), [&](i64 a, i64 b) { return key[a] < key[b]; });
    return perm;
}

// ------------------------------------------------------------ results
// FP64 operation count of clip(subject, cp).area() for one pair, in the
// formulation the GPU kernel runs (each half-space value computed once per
// vertex): +, -, *, /, sqrt, atan2 one operation each.  Not part of the
// reference; it prices the clip kernel against the FP64 roofline.
struct OpCount {
    double flops = 0.0;
    i64 pairs = 0, edges = 0, intersections = 0, fan_terms = 0;
};
constexpr double kOpsCross = 9, kOpsDot = 5, kOpsNrm = 10;
double count_clip_area_ops(const Poly &subject, const Poly &cp, OpCount &oc) {
    double f = 0.0;
    if (subject.n >= 3 && cp.n >= 3) {
        Poly a = subject, b;
        for (int ci = 0; ci < cp.n; ++ci) {
            const V3 &ea = cp.v[ci];
            const V3 &eb = cp.v[(ci + 1) % cp.n];
            V3 normal = cross(ea, eb);
            f += kOpsCross + 5;  // cross, lsq
            if (lsq(normal) < 1.0e-30) continue;
            ++oc.edges;
            b.n = 0;
            constexpr double eps = -1.0e-14;
            f += kOpsDot;  // the first predecessor's value
            for (int i = 0; i < a.n; ++i) {
                int q = (i + a.n - 1) % a.n;
                const V3 &cur = a.v[i];
                const V3 &prv = a.v[q];
                double dc = dot(normal, cur), dp = dot(normal, prv);
                f += kOpsDot;
                bool ci_in = dc >= eps, pi_in = dp >= eps;
                auto inter = [&]() {
                    ++oc.intersections;
                    V3 d = cross(cross(ea, eb), cross(prv, cur));
                    f += 3 * kOpsCross + 5;
                    f += lsq(d) < 1.0e-30 ? 3 + kOpsNrm : kOpsNrm + 3 + kOpsNrm + kOpsDot;
                    return gc_intersection(prv, cur, ea, eb);
                };
                if (ci_in) {
                    if (!pi_in) b.push(inter());
                    b.push(cur);
                } else if (pi_in) {
                    b.push(inter());
                }
            }
            a.n = b.n;
            for (int i = 0; i < b.n; ++i) a.v[i] = b.v[i];
            if (a.n < 3) {
                a.n = 0;
                break;
            }
        }
        if (a.n >= 3) {
            oc.fan_terms += a.n - 2;
            f += 35.0 * (a.n - 2);  // 4 dots, a cross, 3 adds, atan2, 2.0 *, the running sum
        }
    }
    ++oc.pairs;
    oc.flops += f;
    return f;
}

struct Result {
    i64 n_src = 0, n_dst = 0;
    std::vector<i64> row_ptr, col;
    std::vector<double> val;
    std::vector<double> frac_a, frac_b, area_a, area_b;
    // diagnostics
    i64 n_candidates = 0, n_cap_rejected = 0, n_clipped = 0;
    OpCount ops;                              // optional FP64 operation count of the clips
    std::vector<i64> cand_offsets, cand_idx;  // optional broad phase dump
    std::string error;
    int path = 0;  // 0 general, 1 regular fast path, 2 rectilinear fast path
};

// to_csr, interpolation_matrix.hpp:180-255: sort (row, col), keep duplicates.
void to_csr(Result &r, const std::vector<i64> &rows, const std::vector<i64> &cols, const std::vector<double> &vals) {
    const i64 nnz = (i64)vals.size();
    std::vector<i64> perm(nnz);
    std::iota(perm.begin(), perm.end(), 0);
    std::sort(perm.begin(), perm.end(), [&](i64 a, i64 b) {
        if (rows[a] != rows[b]) return rows[a] < rows[b];
        return cols[a] < cols[b];
    });
    r.row_ptr.assign(r.n_dst + 1, 0);
    r.col.resize(nnz);
    r.val.resize(nnz);
    for (i64 k = 0; k < nnz; ++k) r.row_ptr[rows[perm[k]] + 1]++;
    for (i64 i = 0; i < r.n_dst; ++i) r.row_ptr[i + 1] += r.row_ptr[i];
    for (i64 k = 0; k < nnz; ++k) {
        r.col[k] = cols[perm[k]];
        r.val[k] = vals[perm[k]];
    }
}

// generate_conservative_impl<3> (great circle),
// weight_generator.cpp:3304-3418 (setup) and :3586-3769 (loop, packing).
void coastal(const Mesh &src, const Mesh &dst, int extrap, std::vector<i64> &rows, std::vector<i64> &cols, std::vector<double> &vals);

void conservative(const Mesh &src, const Mesh &dst, int norm_type, int unmapped_error, int brute, int dump_candidates, int extrap, Result &r,
                  int parallel_clip, int count_ops) {
    const i64 n_src = src.n_cells, n_dst = dst.n_cells;
    r.n_src = n_src;
    r.n_dst = n_dst;
    RegularInfo sg = detect_regular(src), dg = detect_regular(dst);
    const bool has_smask = src.mask != nullptr, has_dmask = dst.mask != nullptr;
    // DegenerateCellHandler::scan x2 (:3327-3338)
    std::vector<char> sdeg(n_src), ddeg(n_dst);
#pragma omp parallel for schedule(static)
    for (i64 i = 0; i < n_src; ++i) sdeg[i] = classify_cell(src, i) != 0;
#pragma omp parallel for schedule(static)
    for (i64 j = 0; j < n_dst; ++j) ddeg[j] = classify_cell(dst, j) != 0;
    // broad phase (:3345-3355)
    std::vector<FBox> sb = aabbs3d(src), db = aabbs3d(dst);
    std::vector<i64> coff(n_dst + 1, 0);
    std::vector<std::vector<i64>> cl(n_dst);
    if (brute) {
#pragma omp parallel for schedule(dynamic, 64)
        for (i64 j = 0; j < n_dst; ++j)
            for (i64 i = 0; i < n_src; ++i)
                if (box_hit(sb[i], db[j])) cl[j].push_back(i);
    } else {
        Bvh bvh;
        bvh.init(sb);
#pragma omp parallel
        {
            std::vector<i64> tmp;
#pragma omp for schedule(dynamic, 256)
            for (i64 j = 0; j < n_dst; ++j) {
                bvh.query(db[j], tmp);
                cl[j] = tmp;
            }
        }
    }
    for (i64 j = 0; j < n_dst; ++j) coff[j + 1] = coff[j] + (i64)cl[j].size();
    r.n_candidates = coff[n_dst];
    if (dump_candidates) {
        r.cand_offsets = coff;
        r.cand_idx.resize(coff[n_dst]);
        for (i64 j = 0; j < n_dst; ++j) std::copy(cl[j].begin(), cl[j].end(), r.cand_idx.begin() + coff[j]);
    }
    // areas (:3372-3374)
    std::vector<double> sa(n_src), da(n_dst);
    if (src.areas)
        std::copy(src.areas, src.areas + n_src, sa.begin());
    else {
#pragma omp parallel for schedule(static)
        for (i64 i = 0; i < n_src; ++i) sa[i] = cell_poly(src, i).area();
    }
    if (dst.areas)
        std::copy(dst.areas, dst.areas + n_dst, da.begin());
    else {
#pragma omp parallel for schedule(static)
        for (i64 j = 0; j < n_dst; ++j) da[j] = cell_poly(dst, j).area();
    }
    std::vector<i64> order = morton_order(dst);  // :3384-3390
    std::vector<V3> scen, dcen;
    std::vector<double> srad, drad;
    cap_data(src, scen, srad);  // :3397-3399
    cap_data(dst, dcen, drad);
    NodeTrig st = build_node_trig(sg), dt = build_node_trig(dg);  // :3407-3414
    std::vector<double> fa(n_src, 0.0), fb(n_dst, 0.0);
    std::vector<i64> rows, cols;
    std::vector<double> vals;
    // a pair the loop below clips (its filters, in its order)
    auto clipped_pair = [&](i64 j, i64 i) {
        if (has_smask && src.mask[i] == 0) return false;
        if (sdeg[i] || sa[i] <= 0.0) return false;
        return !cap_rejects(scen[i], srad[i], dcen[j], drad[j]);
    };
    auto row_clipped = [&](i64 j) { return !(has_dmask && dst.mask[j] == 0) && !ddeg[j] && da[j] > 0.0; };
    // all-parallel variant: every row's overlaps computed in parallel first;
    // the loop below then uses them in its own order (same results)
    std::vector<std::vector<double>> pre;
    if (parallel_clip) {
        pre.resize(n_dst);
#pragma omp parallel for schedule(dynamic, 64)
        for (i64 j = 0; j < n_dst; ++j) {
            if (!row_clipped(j)) continue;
            pre[j].assign(cl[j].size(), 0.0);
            Poly dp = dt.valid ? cell_poly_cached(dt, j) : cell_poly(dst, j);
            for (size_t q = 0; q < cl[j].size(); ++q) {
                const i64 i = cl[j][q];
                if (!clipped_pair(j, i)) continue;
                Poly sp = st.valid ? cell_poly_cached(st, i) : cell_poly(src, i);
                pre[j][q] = clip(sp, dp).area();
            }
        }
    }
    // sequential overlap loop in Morton order (:3610-3700)
    for (i64 k = 0; k < n_dst; ++k) {
        i64 j = order[k];
        if (has_dmask && dst.mask[j] == 0) continue;
        if (ddeg[j]) continue;
        double adst = da[j];
        if (adst <= 0.0) {
            if (unmapped_error) {
                r.error = "unmapped destination cell " + std::to_string(j);
                return;
            }
            continue;
        }
        bool has_entry = false;
        Poly dp = dt.valid ? cell_poly_cached(dt, j) : cell_poly(dst, j);
        for (size_t q = 0; q < cl[j].size(); ++q) {
            const i64 i = cl[j][q];
            if (has_smask && src.mask[i] == 0) continue;
            if (sdeg[i]) continue;
            double asrc = sa[i];
            if (asrc <= 0.0) continue;
            if (cap_rejects(scen[i], srad[i], dcen[j], drad[j])) {
                ++r.n_cap_rejected;
                continue;
            }
            ++r.n_clipped;
            double ov;
            if (parallel_clip) {
                ov = pre[j][q];
            } else {
                Poly sp = st.valid ? cell_poly_cached(st, i) : cell_poly(src, i);
                ov = clip(sp, dp).area();
                if (count_ops) count_clip_area_ops(sp, dp, r.ops);
            }
            if (ov <= 0.0) continue;
            double w = ov / adst;
            w = std::max(w, 0.0);
            vals.push_back(w);
            rows.push_back(j);
            cols.push_back(i);
            has_entry = true;
            fa[i] += ov / asrc;
            fb[j] += ov / adst;
        }
        if (!has_entry && unmapped_error) {
            r.error = "unmapped destination cell " + std::to_string(j);
            return;
        }
    }
    if (norm_type == 1) {  // FracArea (:3709-3716)
        for (size_t k = 0; k < vals.size(); ++k)
            if (fb[rows[k]] > 0.0) vals[k] /= fb[rows[k]];
    }
    for (auto &x : fa) x = std::min(x, 1.0);  // :3719-3724
    for (auto &x : fb) x = std::min(x, 1.0);
    r.frac_a = fa;
    r.frac_b = fb;
    r.area_a = sa;
    r.area_b = da;
    coastal(src, dst, extrap, rows, cols, vals);
    to_csr(r, rows, cols, vals);
}

// compute_cell_centroids_xy, weight_generator.cpp:74-138: on a sphere the
// mean of the vertices' unit vectors, back to lon in [0, 2pi) and lat in the
// mesh's unit; otherwise the mean of the first two coordinates.
void centroid_lonlat(const Mesh &m, i64 c, double &cx, double &cy) {
    const double pi = 3.14159265358979323846;
    double sx = 0.0, sy = 0.0, sz = 0.0;
    const i64 b = m.offsets[c], e = m.offsets[c + 1];
    if (m.coord_sys == 2) {
        for (i64 k = b; k < e; ++k) {
            sx += m.lon(m.indices[k]);
            sy += m.lat(m.indices[k]);
        }
        const double inv = (e > b) ? 1.0 / (double)(e - b) : 0.0;
        cx = sx * inv;
        cy = sy * inv;
        return;
    }
    for (i64 k = b; k < e; ++k) {
        double lon = m.lon(m.indices[k]), lat = m.lat(m.indices[k]);
        if (m.coord_sys == 0) {
            lon = lon * pi / 180.0;
            lat = lat * pi / 180.0;
        }
        sx += std::cos(lat) * std::cos(lon);
        sy += std::cos(lat) * std::sin(lon);
        sz += std::sin(lat);
    }
    const double inv = (e > b) ? 1.0 / (double)(e - b) : 0.0;
    const double ax = sx * inv, ay = sy * inv, az = sz * inv;
    double lat_avg = std::asin(az);
    double lon_avg = std::atan2(ay, ax);
    if (lon_avg < 0.0) lon_avg += 2.0 * pi;
    if (m.coord_sys == 0) {
        lon_avg = lon_avg * 180.0 / pi;
        lat_avg = lat_avg * 180.0 / pi;
    }
    cx = lon_avg;
    cy = lat_avg;
}

// coastal_renormalize_and_extrapolate, weight_generator.cpp:1736-1881, on
// the generator's COO triplets in their generation order (before to_csr).
// Wet weights (mask > 0) are kept, a row's kept weights divided by their sum
// when that sum is in (0, 1); with NearestWet, a row left without weights
// gets weight 1 at the wet (mask != 0) source cell whose float (lon, lat)
// centroid is nearest in the plane (ArborX nearest on Point<2>; the
// restatement takes the smallest squared float distance, then the smallest
// index -- exact ties are the one case ArborX leaves open).
void coastal(const Mesh &src, const Mesh &dst, int extrap, std::vector<i64> &rows, std::vector<i64> &cols, std::vector<double> &vals) {
    if (src.mask == nullptr) return;
    const i64 n_dst = dst.n_cells, n_src = src.n_cells;
    std::vector<double> row_sums(n_dst, 0.0);
    for (size_t k = 0; k < vals.size(); ++k)
        if (src.mask[cols[k]] > 0) row_sums[rows[k]] += vals[k];
    std::vector<i64> ur, uc;
    std::vector<double> uv;
    for (size_t k = 0; k < vals.size(); ++k) {
        if (src.mask[cols[k]] > 0) {
            double w = vals[k];
            const double s = row_sums[rows[k]];
            if (s > 0.0 && s < 1.0) w /= s;
            ur.push_back(rows[k]);
            uc.push_back(cols[k]);
            uv.push_back(w);
        }
    }
    std::vector<char> has(n_dst, 0);
    for (i64 r : ur) has[r] = 1;
    if (extrap == 1) {
        std::vector<float> wx, wy;
        std::vector<i64> wi;
        for (i64 i = 0; i < n_src; ++i) {
            if (src.mask[i] != 0) {
                double cx, cy;
                centroid_lonlat(src, i, cx, cy);
                wx.push_back((float)cx);
                wy.push_back((float)cy);
                wi.push_back(i);
            }
        }
        if (!wi.empty()) {
            for (i64 j = 0; j < n_dst; ++j) {
                if (has[j]) continue;
                double cx, cy;
                centroid_lonlat(dst, j, cx, cy);
                const float qx = (float)cx, qy = (float)cy;
                float best = std::numeric_limits<float>::infinity();
                i64 bi = -1;
                for (size_t q = 0; q < wi.size(); ++q) {
                    const float dx = wx[q] - qx, dy = wy[q] - qy;
                    const float d2 = dx * dx + dy * dy;
                    if (d2 < best || (d2 == best && bi < 0)) {
                        best = d2;
                        bi = (i64)q;
                    }
                }
                if (bi < 0) continue;  // NaN distances only
                ur.push_back(j);
                uc.push_back(wi[bi]);
                uv.push_back(1.0);
            }
        }
    }
    rows.swap(ur);
    cols.swap(uc);
    vals.swap(uv);
}

// ------------------------------------------- regular / rectilinear fast paths
// rect_overlap_spherical (weight_generator_conservative_rect.cpp:51-69, and
// its twin in weight_generator_conservative_rect_nonuniform.cpp:23-40):
// longitude overlap times the difference of the sines of the clamped
// latitudes, degrees converted with the folded constant pi / 180.
double rect_overlap_sph(double slx, double shx, double sly, double shy, double dlx, double dhx, double dly, double dhy, bool deg) {
    double dx = std::fmax(0.0, std::fmin(shx, dhx) - std::fmax(slx, dlx));
    double lo = std::fmax(sly, dly), hi = std::fmin(shy, dhy);
    if (hi <= lo || dx <= 0.0) return 0.0;
    if (deg) {
        const double d2r = 3.14159265358979323846 / 180.0;
        dx *= d2r;
        lo *= d2r;
        hi *= d2r;
    }
    return dx * (std::sin(hi) - std::sin(lo));
}
double rect_overlap_flat(double slx, double shx, double sly, double shy, double dlx, double dhx, double dly, double dhy) {
    const double dx = std::fmax(0.0, std::fmin(shx, dhx) - std::fmax(slx, dlx));
    const double dy = std::fmax(0.0, std::fmin(shy, dhy) - std::fmax(sly, dly));
    return dx * dy;
}
// Band area dlon * (sin(hi) - sin(lo)) (or dlon * dlat in the plane).
double band_area(double dlon, double lo, double hi, int cs) {
    if (cs == 2) return dlon * (hi - lo);
    if (cs == 0) {
        const double d2r = 3.14159265358979323846 / 180.0;
        dlon *= d2r;
        lo *= d2r;
        hi *= d2r;
    }
    return dlon * (std::sin(hi) - std::sin(lo));
}

// generate_conservative_rect, weight_generator_conservative_rect.cpp:87-395.
// Returns false where the reference returns its empty "decline" matrix.
bool conservative_regular(const Mesh &src, const Mesh &dst, const RegularInfo &sg, const RegularInfo &dg, int norm_type, std::vector<i64> &rows,
                          std::vector<i64> &cols, std::vector<double> &vals, Result &r) {
    const i64 n_src = src.n_cells, n_dst = dst.n_cells;
    const bool deg = src.coord_sys == 0, sph = src.coord_sys != 2;
    const double wrap = deg ? 350.0 : 6.0;
    const bool s_glob = (double)sg.ni * sg.delta_lon > wrap, d_glob = (double)dg.ni * dg.delta_lon > wrap;
    if (!s_glob && !d_glob && (dg.lon_min < sg.lon_min - sg.delta_lon * 0.5 || dg.lon_max > sg.lon_max + sg.delta_lon * 0.5)) return false;
    std::vector<double> sa(n_src), da(n_dst), fa(n_src, 0.0), fb(n_dst, 0.0);
    for (i64 c = 0; c < n_src; ++c) {
        if (src.areas) {
            sa[c] = src.areas[c];
        } else {
            const double lo = sg.lat_min + (double)(c / sg.ni) * sg.delta_lat;
            sa[c] = sph ? band_area(sg.delta_lon, lo, lo + sg.delta_lat, src.coord_sys) : sg.delta_lon * sg.delta_lat;
        }
    }
    for (i64 c = 0; c < n_dst; ++c) {
        if (dst.areas) {
            da[c] = dst.areas[c];
        } else {
            const double lo = dg.lat_min + (double)(c / dg.ni) * dg.delta_lat;
            da[c] = sph ? band_area(dg.delta_lon, lo, lo + dg.delta_lat, src.coord_sys) : dg.delta_lon * dg.delta_lat;
        }
    }
    for (i64 jd = 0; jd < dg.nj; ++jd) {
        for (i64 id = 0; id < dg.ni; ++id) {
            const i64 j = jd * dg.ni + id;
            if (dst.mask && dst.mask[j] == 0) continue;
            const double dlx = dg.lon_min + (double)id * dg.delta_lon, dhx = dlx + dg.delta_lon;
            const double dly = dg.lat_min + (double)jd * dg.delta_lat, dhy = dly + dg.delta_lat;
            double nlx = dlx, nhx = dhx;
            if (s_glob) {
                const double shift = 360.0 * std::floor((dlx - sg.lon_min) / 360.0);
                nlx = dlx - shift;
                nhx = dhx - shift;
            }
            const double ad = da[j];
            if (ad <= 0.0) continue;
            int is_lo = (int)std::floor((nlx - sg.lon_min) / sg.delta_lon);
            int is_hi = (int)std::floor((nhx - sg.lon_min) / sg.delta_lon);
            int js_lo = (int)std::floor((dly - sg.lat_min) / sg.delta_lat);
            int js_hi = (int)std::floor((dhy - sg.lat_min) / sg.delta_lat);
            if (!s_glob) {
                is_lo = std::max(0, is_lo);
                is_hi = std::min((int)sg.ni - 1, is_hi);
            }
            js_lo = std::max(0, js_lo);
            js_hi = std::min((int)sg.nj - 1, js_hi);
            const bool wraps = s_glob && is_lo > is_hi;
            if (!wraps && is_lo > is_hi) continue;
            if (js_lo > js_hi) continue;
            for (int js = js_lo; js <= js_hi; ++js) {
                for (int is = is_lo; is <= is_hi; ++is) {
                    int ia = is;
                    if (s_glob) ia = ((is % (int)sg.ni) + (int)sg.ni) % (int)sg.ni;
                    const i64 i = (i64)js * sg.ni + ia;
                    if (i >= n_src) continue;
                    if (src.mask && src.mask[i] == 0) continue;
                    const double as = sa[i];
                    if (as <= 0.0) continue;
                    const double slx = sg.lon_min + (double)is * sg.delta_lon, shx = slx + sg.delta_lon;
                    const double sly = sg.lat_min + (double)js * sg.delta_lat, shy = sly + sg.delta_lat;
                    const double ov = sph ? rect_overlap_sph(slx, shx, sly, shy, nlx, nhx, dly, dhy, deg) : rect_overlap_flat(slx, shx, sly, shy, nlx, nhx, dly, dhy);
                    if (ov <= 0.0) continue;
                    vals.push_back(std::max(ov / ad, 0.0));
                    rows.push_back(j);
                    cols.push_back(i);
                    fa[i] += ov / as;
                    fb[j] += ov / ad;
                }
            }
        }
    }
    if (norm_type == 1)
        for (size_t k = 0; k < vals.size(); ++k)
            if (fb[rows[k]] > 0.0) vals[k] /= fb[rows[k]];
    for (auto &x : fa) x = std::min(x, 1.0);
    for (auto &x : fb) x = std::min(x, 1.0);
    r.frac_a = fa;
    r.frac_b = fb;
    r.area_a = sa;
    r.area_b = da;
    return true;
}

// generate_conservative_rect_nonuniform,
// weight_generator_conservative_rect_nonuniform.cpp:52-313.
void conservative_rectilinear(const Mesh &src, const Mesh &dst, const RectInfo &sr, const RectInfo &dr, int norm_type, std::vector<i64> &rows,
                              std::vector<i64> &cols, std::vector<double> &vals, Result &r) {
    const i64 n_src = src.n_cells, n_dst = dst.n_cells, sni = sr.ni, snj = sr.nj, dni = dr.ni, dnj = dr.nj;
    const bool deg = src.coord_sys == 0, sph = src.coord_sys != 2;
    const double slon0 = sr.ulon[0], slon1 = sr.ulon[sni];
    const bool s_glob = slon1 - slon0 > (deg ? 350.0 : 6.0);
    std::vector<double> fa(n_src, 0.0), fb(n_dst, 0.0);
    auto area = [&](double lx, double hx, double ly, double hy) { return sph ? band_area(hx - lx, ly, hy, src.coord_sys) : (hx - lx) * (hy - ly); };
    for (i64 jd = 0; jd < dnj; ++jd) {
        for (i64 id = 0; id < dni; ++id) {
            const i64 j = id + jd * dni;
            if (dst.mask && dst.mask[j] == 0) continue;
            const double dlx = dr.ulon[id], dhx = dr.ulon[id + 1], dly = dr.ulat[jd], dhy = dr.ulat[jd + 1];
            double nlx = dlx, nhx = dhx;
            if (s_glob) {
                const double shift = 360.0 * std::floor((dlx - slon0) / 360.0);
                nlx = dlx - shift;
                nhx = dhx - shift;
            }
            const double ad = dst.areas ? dst.areas[j] : area(nlx, nhx, dly, dhy);
            if (ad <= 0.0) continue;
            const i64 a0 = std::lower_bound(sr.ulon.begin(), sr.ulon.end(), nlx) - sr.ulon.begin();
            const i64 a1 = std::upper_bound(sr.ulon.begin(), sr.ulon.end(), nhx) - sr.ulon.begin();
            int is0 = std::max(0, (int)(a0 - 1)), is1 = (int)a1;
            if (s_glob) {
                if (nhx <= slon1) is1 = std::min((int)sni - 1, is1);
            } else {
                is1 = std::min((int)sni - 1, is1);
            }
            const i64 b0 = std::lower_bound(sr.ulat.begin(), sr.ulat.end(), dly) - sr.ulat.begin();
            const i64 b1 = std::upper_bound(sr.ulat.begin(), sr.ulat.end(), dhy) - sr.ulat.begin();
            const int js0 = std::max(0, (int)(b0 - 1)), js1 = std::min((int)snj - 1, (int)b1);
            for (int js = js0; js <= js1; ++js) {
                for (int is = is0; is <= is1; ++is) {
                    const int ia = s_glob ? ((is % (int)sni) + (int)sni) % (int)sni : is;
                    const i64 i = ia + (i64)js * sni;
                    if (src.mask && src.mask[i] == 0) continue;
                    double slx = sr.ulon[ia], shx = sr.ulon[ia + 1];
                    if (s_glob && is >= (int)sni) {
                        slx += 360.0;
                        shx += 360.0;
                    }
                    const double sly = sr.ulat[js], shy = sr.ulat[js + 1];
                    const double ov = sph ? rect_overlap_sph(slx, shx, sly, shy, nlx, nhx, dly, dhy, deg) : rect_overlap_flat(slx, shx, sly, shy, nlx, nhx, dly, dhy);
                    if (ov > 1e-12) {
                        const double as = src.areas ? src.areas[i] : area(slx, shx, sly, shy);
                        if (as <= 0.0) continue;
                        vals.push_back(ov / ad);
                        rows.push_back(j);
                        cols.push_back(i);
                        fa[i] += ov / as;
                        fb[j] += ov / ad;
                    }
                }
            }
        }
    }
    if (norm_type == 1)
        for (size_t k = 0; k < vals.size(); ++k)
            if (fb[rows[k]] > 0.0) vals[k] /= fb[rows[k]];
    r.frac_a.resize(n_src);
    r.area_a.resize(n_src);
    r.frac_b.resize(n_dst);
    r.area_b.resize(n_dst);
    for (i64 i = 0; i < n_src; ++i) {
        r.area_a[i] = src.areas ? src.areas[i] : area(sr.ulon[i % sni], sr.ulon[i % sni + 1], sr.ulat[i / sni], sr.ulat[i / sni + 1]);
        r.frac_a[i] = std::min(fa[i], 1.0);
    }
    for (i64 j = 0; j < n_dst; ++j) {
        r.area_b[j] = dst.areas ? dst.areas[j] : area(dr.ulon[j % dni], dr.ulon[j % dni + 1], dr.ulat[j / dni], dr.ulat[j / dni + 1]);
        r.frac_b[j] = std::min(fb[j], 1.0);
    }
}

// generate_bilinear_rect, weight_generator_bilinear_rect.cpp:33-266: the
// planar mean of the cell's nodes, fractional cell-centred position, 2x2
// stencil, longitude periodic when the grid spans 360 units.  Returns false
// where the reference declines (ni or nj zero).
bool bilinear_regular(const Mesh &src, const Mesh &dst, const RegularInfo &sg, int unmapped_error, std::vector<i64> &rows, std::vector<i64> &cols,
                      std::vector<double> &vals, Result &r) {
    const i64 ni = sg.ni, nj = sg.nj;
    if (ni == 0 || nj == 0) return false;
    const bool per = std::abs((sg.lon_max - sg.lon_min) - 360.0) < 1.0e-10 * sg.delta_lon;
    r.frac_a.assign(src.n_cells, 1.0);
    r.frac_b.assign(dst.n_cells, 0.0);
    r.area_a.assign(src.n_cells, 0.0);
    r.area_b.assign(dst.n_cells, 0.0);
    for (i64 c = 0; c < dst.n_cells; ++c) {
        const i64 b = dst.offsets[c], e = dst.offsets[c + 1];
        double sx = 0.0, sy = 0.0;
        for (i64 k = b; k < e; ++k) {
            sx += dst.lon(dst.indices[k]);
            sy += dst.lat(dst.indices[k]);
        }
        double lon = sx / (double)(e - b), lat = sy / (double)(e - b);
        if (per) {
            double off = lon - sg.lon_min;
            off = off - std::floor(off / 360.0) * 360.0;
            lon = sg.lon_min + off;
        }
        const double fi = (lon - sg.lon_min) / sg.delta_lon - 0.5, fj = (lat - sg.lat_min) / sg.delta_lat - 0.5;
        const double tol = 1.0e-9;
        bool out = fj < -0.5 - tol || fj > (double)nj - 0.5 + tol;
        if (!per) out = out || fi < -0.5 - tol || fi > (double)ni - 0.5 + tol;
        if (out) {
            if (unmapped_error) {
                r.error = "Unmapped destination cell " + std::to_string(c) + " in bilinear regular-grid fast-path";
                return true;
            }
            continue;
        }
        int i = (int)std::floor(fi), j = (int)std::floor(fj);
        if (j < 0) j = 0;
        if (j >= (int)nj - 1) j = (int)nj - 2;
        if (j < 0) j = 0;
        if (!per) {
            if (i < 0) i = 0;
            if (i >= (int)ni - 1) i = (int)ni - 2;
            if (i < 0) i = 0;
        } else {
            i = ((i % (int)ni) + (int)ni) % (int)ni;
        }
        const double tx = std::min(std::max(fi - (double)i, 0.0), 1.0), ty = std::min(std::max(fj - (double)j, 0.0), 1.0);
        int i1 = i + 1, j1 = j + 1;
        if (i1 >= (int)ni) i1 = per ? 0 : (int)ni - 1;
        if (j1 >= (int)nj) j1 = (int)nj - 1;
        const i64 ix[4] = {(i64)j * ni + i, (i64)j * ni + i1, (i64)j1 * ni + i, (i64)j1 * ni + i1};
        const double w[4] = {(1.0 - tx) * (1.0 - ty), tx * (1.0 - ty), (1.0 - tx) * ty, tx * ty};
        for (int k = 0; k < 4; ++k) {
            rows.push_back(c);
            cols.push_back(ix[k]);
            vals.push_back(w[k]);
        }
        r.frac_b[c] = 1.0;
    }
    return true;
}

// get_cell_areas / compute_single_cell_area, weight_generator.cpp:264-313:
// the mesh's areas, else the shoelace area in raw coordinates.
std::vector<double> planar_areas(const Mesh &m) {
    std::vector<double> a(m.n_cells);
    for (i64 c = 0; c < m.n_cells; ++c) {
        if (m.areas) {
            a[c] = m.areas[c];
            continue;
        }
        const i64 b = m.offsets[c], n = m.offsets[c + 1] - b;
        if (n < 3) {
            a[c] = 0.0;
            continue;
        }
        double s = 0.0;
        for (i64 i = 0; i < n; ++i) {
            const i64 p = m.indices[b + i], q = m.indices[b + (i + 1) % n];
            s += m.lon(p) * m.lat(q) - m.lon(q) * m.lat(p);
        }
        a[c] = std::abs(s) * 0.5;
    }
    return a;
}

// generate_nearest_rect, weight_generator.cpp:2025-2083: the dst cell's
// (lon_min + i dlon, lat_min + j dlat) rounded to the nearest source index,
// longitude wrapped, latitude clamped; fractions and areas left zero.
void nearest_rect(const Mesh &src, const Mesh &dst, const RegularInfo &sg, const RegularInfo &dg, std::vector<i64> &rows, std::vector<i64> &cols,
                  std::vector<double> &vals, Result &r) {
    for (i64 j = 0; j < dst.n_cells; ++j) {
        const i64 di = j % dg.ni, dj = j / dg.ni;
        const double lon = dg.lon_min + (double)di * dg.delta_lon;
        const double lat = dg.lat_min + (double)dj * dg.delta_lat;
        i64 si = (i64)std::round((lon - sg.lon_min) / sg.delta_lon);
        si = si % sg.ni;
        if (si < 0) si += sg.ni;
        i64 sj = (i64)std::round((lat - sg.lat_min) / sg.delta_lat);
        if (sj < 0) sj = 0;
        if (sj >= sg.nj) sj = sg.nj - 1;
        rows.push_back(j);
        cols.push_back(sj * sg.ni + si);
        vals.push_back(1.0);
    }
    r.frac_a.assign(src.n_cells, 0.0);
    r.frac_b.assign(dst.n_cells, 0.0);
    r.area_a.assign(src.n_cells, 0.0);
    r.area_b.assign(dst.n_cells, 0.0);
}

// generate_nearest_impl<3> / <2>, weight_generator.cpp:2129-2271 (host
// path): float points from the double centroids -- unit vectors
// (cos lat cos lon, cos lat sin lon, sin lat) on a sphere, (lon, lat)
// otherwise -- and each destination's nearest source point by float squared
// distance (brute force here; ties to the smallest index, which ArborX leaves
// open).  Fractions 1, areas from get_cell_areas.
void nearest_points(const Mesh &src, const Mesh &dst, int unmapped_error, std::vector<i64> &rows, std::vector<i64> &cols,
                    std::vector<double> &vals, Result &r) {
    const double pi = 3.14159265358979323846;
    const bool sph = src.coord_sys != 2;
    auto points = [&](const Mesh &m) {
        std::vector<float> p(3 * m.n_cells);
        for (i64 c = 0; c < m.n_cells; ++c) {
            double lon, lat;
            centroid_lonlat(m, c, lon, lat);
            if (sph) {
                if (m.coord_sys == 0) {
                    lon = lon * pi / 180.0;
                    lat = lat * pi / 180.0;
                }
                p[3 * c] = (float)(std::cos(lat) * std::cos(lon));
                p[3 * c + 1] = (float)(std::cos(lat) * std::sin(lon));
                p[3 * c + 2] = (float)std::sin(lat);
            } else {
                p[3 * c] = (float)lon;
                p[3 * c + 1] = (float)lat;
                p[3 * c + 2] = 0.0f;
            }
        }
        return p;
    };
    const std::vector<float> sp = points(src), dp = points(dst);
    std::vector<i64> best(dst.n_cells, -1);
#pragma omp parallel for schedule(dynamic, 64)
    for (i64 j = 0; j < dst.n_cells; ++j) {
        float bd = std::numeric_limits<float>::infinity();
        i64 bi = -1;
        for (i64 i = 0; i < src.n_cells; ++i) {
            const float dx = sp[3 * i] - dp[3 * j], dy = sp[3 * i + 1] - dp[3 * j + 1], dz = sp[3 * i + 2] - dp[3 * j + 2];
            const float d2 = sph ? dx * dx + dy * dy + dz * dz : dx * dx + dy * dy;
            if (d2 < bd || (d2 == bd && bi < 0)) {
                bd = d2;
                bi = i;
            }
        }
        best[j] = bi;
    }
    for (i64 j = 0; j < dst.n_cells; ++j) {
        if (best[j] < 0) {
            if (unmapped_error) {
                r.error = "WeightGenerator::generate_nearest: unmapped destination cell " + std::to_string(j);
                return;
            }
            continue;
        }
        rows.push_back(j);
        cols.push_back(best[j]);
        vals.push_back(1.0);
    }
    r.frac_a.assign(src.n_cells, 1.0);
    r.frac_b.assign(dst.n_cells, 1.0);
    r.area_a = planar_areas(src);
    r.area_b = planar_areas(dst);
}

// generate_bilinear_rect_nonuniform, weight_generator_bilinear_rect_nonuniform.cpp:22-149
void bilinear_rect(const Mesh &src, const Mesh &dst, const RectInfo &ri, Result &r) {
    const i64 ni = ri.ni, nj = ri.nj, n_dst = dst.n_cells;
    r.n_src = src.n_cells;
    r.n_dst = n_dst;
    std::vector<double> cl(ni), ca(nj);
    for (i64 i = 0; i < ni; ++i) cl[i] = 0.5 * (ri.ulon[i] + ri.ulon[i + 1]);
    for (i64 j = 0; j < nj; ++j) ca[j] = 0.5 * (ri.ulat[j] + ri.ulat[j + 1]);
    r.row_ptr.resize(n_dst + 1);
    r.col.resize(4 * n_dst);
    r.val.resize(4 * n_dst);
    for (i64 c = 0; c < n_dst; ++c) {
        i64 b = dst.offsets[c], e = dst.offsets[c + 1];
        double ls = 0, as = 0;
        for (i64 k = b; k < e; ++k) {
            ls += dst.lon(dst.indices[k]);
            as += dst.lat(dst.indices[k]);
        }
        double lon = ls / (double)(e - b), lat = as / (double)(e - b);
        int i = (int)(std::upper_bound(cl.begin(), cl.end(), lon) - cl.begin()) - 1;
        int j = (int)(std::upper_bound(ca.begin(), ca.end(), lat) - ca.begin()) - 1;
        if (i < 0) i = 0;
        if (i >= (int)ni - 1) i = (int)ni - 2;
        if (i < 0) i = 0;
        if (j < 0) j = 0;
        if (j >= (int)nj - 1) j = (int)nj - 2;
        if (j < 0) j = 0;
        double x0 = cl[i], x1 = cl[i + 1], y0 = ca[j], y1 = ca[j + 1];
        double tx = (lon - x0) / (x1 - x0), ty = (lat - y0) / (y1 - y0);
        tx = std::max(0.0, std::min(1.0, tx));
        ty = std::max(0.0, std::min(1.0, ty));
        i64 i1 = i + 1, j1 = j + 1;
        i64 idx[4] = {(i64)j * ni + i, (i64)j * ni + i1, j1 * ni + i, j1 * ni + i1};
        double w[4] = {(1.0 - tx) * (1.0 - ty), tx * (1.0 - ty), (1.0 - tx) * ty, tx * ty};
        // to_csr's (row, col) sort: the four columns are already ascending
        // whenever ni >= 2 (i < i1, j < j1); sort defensively otherwise.
        int ord[4] = {0, 1, 2, 3};
        std::stable_sort(ord, ord + 4, [&](int p, int q) { return idx[p] < idx[q]; });
        r.row_ptr[c] = 4 * c;
        for (int k = 0; k < 4; ++k) {
            r.col[4 * c + k] = idx[ord[k]];
            r.val[4 * c + k] = w[ord[k]];
        }
    }
    r.row_ptr[n_dst] = 4 * n_dst;
    r.frac_a.assign(src.n_cells, 1.0);
    r.frac_b.assign(n_dst, 1.0);
    r.area_a.assign(src.n_cells, 0.0);
    r.area_b.assign(n_dst, 0.0);
}

// WeightGenerator::generate checks, weight_generator.cpp:1887-1992.
std::string generate_checks(const Mesh &src, const Mesh &dst, int method) {
    if (src.coord_sys != dst.coord_sys) return "coordinate system mismatch";
    if (src.coord_sys == 0 || src.coord_sys == 1) {
        auto range = [](const Mesh &m, double &lo, double &hi) {
            lo = std::numeric_limits<double>::max();
            hi = -std::numeric_limits<double>::max();
            for (i64 c = 0; c < m.n_cells; ++c) {
                i64 b = m.offsets[c], e = m.offsets[c + 1];
                if (e == b) continue;
                double s = 0;
                for (i64 k = b; k < e; ++k) s += m.lon(m.indices[k]);
                double cx = s / (double)(e - b);
                lo = std::min(lo, cx);
                hi = std::max(hi, cx);
            }
        };
        double s0, s1, d0, d1;
        range(src, s0, s1);
        range(dst, d0, d1);
        if (s1 < d0 || d1 < s0) {
            double period = src.coord_sys == 0 ? 360.0 : 2.0 * kPi;
            double shift = std::round((0.5 * (s0 + s1) - 0.5 * (d0 + d1)) / period) * period;
            bool overlaps = !(s1 < d0 + shift || d1 + shift < s0);
            bool structured = false;
            if (method == 0) structured = detect_regular(src).is_regular || detect_rect(src).is_rect;
            if (overlaps && !structured) return "spherical longitude coordinate range mismatch";
        }
    }
    return "";
}

Mesh make_mesh(const double *coords, i64 n_nodes, const i64 *offsets, const i64 *indices, i64 n_cells, int cs, const double *areas,
               const int32_t *mask) {
    Mesh m;
    m.coords = coords;
    m.n_nodes = n_nodes;
    m.offsets = offsets;
    m.indices = indices;
    m.n_cells = n_cells;
    m.coord_sys = cs;
    m.areas = areas;
    m.mask = mask;
    return m;
}

}  // namespace

extern "C" {

// ---------------------------------------------------------- primitives
double om_polygon_area(const double *xyz, int n) {
    Poly p;
    for (int i = 0; i < n; ++i) p.push(V3{xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]});
    return p.area();
}
int om_clip(const double *subj, int ns, const double *cp, int nc, double *out, int *nout) {
    Poly a, b;
    for (int i = 0; i < ns; ++i) a.push(V3{subj[3 * i], subj[3 * i + 1], subj[3 * i + 2]});
    for (int i = 0; i < nc; ++i) b.push(V3{cp[3 * i], cp[3 * i + 1], cp[3 * i + 2]});
    Poly r = clip(a, b);
    for (int i = 0; i < r.n; ++i) {
        out[3 * i] = r.v[i].x;
        out[3 * i + 1] = r.v[i].y;
        out[3 * i + 2] = r.v[i].z;
    }
    *nout = r.n;
    return 0;
}
double om_overlap_area(const double *subj, int ns, const double *cp, int nc) {
    Poly a, b;
    for (int i = 0; i < ns; ++i) a.push(V3{subj[3 * i], subj[3 * i + 1], subj[3 * i + 2]});
    for (int i = 0; i < nc; ++i) b.push(V3{cp[3 * i], cp[3 * i + 1], cp[3 * i + 2]});
    return clip(a, b).area();
}
void om_lonlat_to_xyz(double lon_rad, double lat_rad, double *out) {
    V3 p = lonlat_xyz(lon_rad, lat_rad);
    out[0] = p.x;
    out[1] = p.y;
    out[2] = p.z;
}

// ------------------------------------------------------- mesh analyses
int om_detect_regular(const double *coords, int64_t n_nodes, const int64_t *offsets, const int64_t *indices, int64_t n_cells, double *out6,
                      int64_t *out_ninj) {
    Mesh m = make_mesh(coords, n_nodes, offsets, indices, n_cells, 0, nullptr, nullptr);
    RegularInfo g = detect_regular(m);
    out6[0] = g.lon_min;
    out6[1] = g.lon_max;
    out6[2] = g.delta_lon;
    out6[3] = g.lat_min;
    out6[4] = g.lat_max;
    out6[5] = g.delta_lat;
    out_ninj[0] = g.ni;
    out_ninj[1] = g.nj;
    return g.is_regular ? 1 : 0;
}
int om_detect_rectilinear(const double *coords, int64_t n_nodes, const int64_t *offsets, const int64_t *indices, int64_t n_cells,
                          int64_t *out_ninj) {
    Mesh m = make_mesh(coords, n_nodes, offsets, indices, n_cells, 0, nullptr, nullptr);
    RectInfo g = detect_rect(m);
    out_ninj[0] = g.ni;
    out_ninj[1] = g.nj;
    return g.is_rect ? 1 : 0;
}
void om_cell_areas(const double *coords, int64_t n_nodes, const int64_t *offsets, const int64_t *indices, int64_t n_cells, int cs,
                   double *out) {
    Mesh m = make_mesh(coords, n_nodes, offsets, indices, n_cells, cs, nullptr, nullptr);
#pragma omp parallel for schedule(static)
    for (i64 c = 0; c < n_cells; ++c) out[c] = cell_poly(m, c).area();
}
void om_degenerate(const double *coords, int64_t n_nodes, const int64_t *offsets, const int64_t *indices, int64_t n_cells, int32_t *out) {
    Mesh m = make_mesh(coords, n_nodes, offsets, indices, n_cells, 0, nullptr, nullptr);
#pragma omp parallel for schedule(static)
    for (i64 c = 0; c < n_cells; ++c) out[c] = classify_cell(m, c);
}
void om_cap_data(const double *coords, int64_t n_nodes, const int64_t *offsets, const int64_t *indices, int64_t n_cells, int cs,
                 double *cen3, double *rad) {
    Mesh m = make_mesh(coords, n_nodes, offsets, indices, n_cells, cs, nullptr, nullptr);
    std::vector<V3> c;
    std::vector<double> r;
    cap_data(m, c, r);
    for (i64 i = 0; i < n_cells; ++i) {
        cen3[3 * i] = c[i].x;
        cen3[3 * i + 1] = c[i].y;
        cen3[3 * i + 2] = c[i].z;
        rad[i] = r[i];
    }
}
int om_cap_rejects(const double *cs, double rs, const double *cd, double rd) {
    return cap_rejects(V3{cs[0], cs[1], cs[2]}, rs, V3{cd[0], cd[1], cd[2]}, rd) ? 1 : 0;
}
// compute_cell_centroids_xy rounded to float (weight_generator.cpp:1813).
void om_cell_centroids(const double *coords, int64_t n_nodes, const int64_t *offsets, const int64_t *indices, int64_t n_cells, int cs, float *lon,
                       float *lat) {
    Mesh m = make_mesh(coords, n_nodes, offsets, indices, n_cells, cs, nullptr, nullptr);
    for (i64 c = 0; c < n_cells; ++c) {
        double cx, cy;
        centroid_lonlat(m, c, cx, cy);
        lon[c] = (float)cx;
        lat[c] = (float)cy;
    }
}
void om_aabbs(const double *coords, int64_t n_nodes, const int64_t *offsets, const int64_t *indices, int64_t n_cells, int cs, float *out6) {
    Mesh m = make_mesh(coords, n_nodes, offsets, indices, n_cells, cs, nullptr, nullptr);
    auto b = aabbs3d(m);
    for (i64 c = 0; c < n_cells; ++c)
        for (int a = 0; a < 3; ++a) {
            out6[6 * c + a] = b[c].lo[a];
            out6[6 * c + 3 + a] = b[c].hi[a];
        }
}

// ------------------------------------------------------- weight generation
// method: 0 Bilinear, 2 Conservative1stOrder (axis_c_interop.cpp:126-140).
// flags bit0: brute-force broad phase; bit1: keep the candidate lists.
void *om_generate(const double *s_coords, int64_t s_nn, const int64_t *s_off, const int64_t *s_idx, int64_t s_nc, const double *s_areas,
                  const int32_t *s_mask, const double *d_coords, int64_t d_nn, const int64_t *d_off, const int64_t *d_idx, int64_t d_nc,
                  const double *d_areas, const int32_t *d_mask, int coord_sys, int method, int norm_type, int unmapped_error, int flags,
                  int extrap, int n_threads) {
    if (n_threads > 0) omp_set_num_threads(n_threads);
    Result *r = new Result;
    Mesh src = make_mesh(s_coords, s_nn, s_off, s_idx, s_nc, coord_sys, s_areas, s_mask);
    Mesh dst = make_mesh(d_coords, d_nn, d_off, d_idx, d_nc, coord_sys, d_areas, d_mask);
    r->error = generate_checks(src, dst, method);
    if (!r->error.empty()) return r;
    r->n_src = src.n_cells;
    r->n_dst = dst.n_cells;
    if (method == 2) {
        // generate_conservative, weight_generator.cpp:3246-3302: the analytic
        // regular, then rectilinear, paths when BOTH meshes qualify; a
        // declined or (under UnmappedAction::Error) empty result falls
        // through to the next path.
        const RegularInfo sg = detect_regular(src), dg = detect_regular(dst);
        if (sg.is_regular && dg.is_regular) {
            std::vector<i64> rows, cols;
            std::vector<double> vals;
            Result t;
            t.n_src = src.n_cells;
            t.n_dst = dst.n_cells;
            if (conservative_regular(src, dst, sg, dg, norm_type, rows, cols, vals, t) && dst.n_cells > 0 && (!unmapped_error || !vals.empty())) {
                *r = t;
                r->path = 1;
                coastal(src, dst, extrap, rows, cols, vals);
                to_csr(*r, rows, cols, vals);
                return r;
            }
        }
        const RectInfo sr = detect_rect(src), dr = detect_rect(dst);
        if (sr.is_rect && dr.is_rect) {
            std::vector<i64> rows, cols;
            std::vector<double> vals;
            Result t;
            t.n_src = src.n_cells;
            t.n_dst = dst.n_cells;
            conservative_rectilinear(src, dst, sr, dr, norm_type, rows, cols, vals, t);
            if (dst.n_cells > 0 && (!unmapped_error || !vals.empty())) {
                *r = t;
                r->path = 2;
                coastal(src, dst, extrap, rows, cols, vals);
                to_csr(*r, rows, cols, vals);
                return r;
            }
        }
        conservative(src, dst, norm_type, unmapped_error, flags & 1, flags & 2, extrap, *r, flags & 4, flags & 8);
    } else if (method == 0) {
        // generate_bilinear, weight_generator.cpp:2306-2345
        const RegularInfo sg = detect_regular(src);
        if (sg.is_regular && detect_regular(dst).is_regular) {
            std::vector<i64> rows, cols;
            std::vector<double> vals;
            if (bilinear_regular(src, dst, sg, unmapped_error, rows, cols, vals, *r)) {
                if (!r->error.empty()) return r;
                r->path = 1;
                coastal(src, dst, extrap, rows, cols, vals);
                to_csr(*r, rows, cols, vals);
                return r;
            }
        }
        RectInfo ri = detect_rect(src);
        if (!ri.is_rect) {
            r->error = "BVH bilinear not restated";
            return r;
        }
        bilinear_rect(src, dst, ri, *r);
        if (src.mask != nullptr) {
            // coastal pass (:2018) over the CSR triplets: the reference's
            // COO lists a row's four corners in another order, which moves
            // the wet-weight sums by rounding only.
            std::vector<i64> rows, cols(r->col.begin(), r->col.end());
            std::vector<double> vals(r->val.begin(), r->val.end());
            for (i64 j = 0; j < dst.n_cells; ++j)
                for (i64 e = r->row_ptr[j]; e < r->row_ptr[j + 1]; ++e) rows.push_back(j);
            coastal(src, dst, extrap, rows, cols, vals);
            to_csr(*r, rows, cols, vals);
        }
    } else if (method == 1) {
        // generate_nearest (host), weight_generator.cpp:2098-2127, then the
        // coastal pass (:2018)
        r->n_src = src.n_cells;
        r->n_dst = dst.n_cells;
        std::vector<i64> rows, cols;
        std::vector<double> vals;
        const RegularInfo sg = detect_regular(src), dg = detect_regular(dst);
        if (sg.is_regular && dg.is_regular)
            nearest_rect(src, dst, sg, dg, rows, cols, vals, *r);
        else
            nearest_points(src, dst, unmapped_error, rows, cols, vals, *r);
        if (!r->error.empty()) return r;
        coastal(src, dst, extrap, rows, cols, vals);
        to_csr(*r, rows, cols, vals);
    } else {
        r->error = "method not restated";
    }
    return r;
}
const char *om_result_error(void *h) { return static_cast<Result *>(h)->error.c_str(); }
int64_t om_result_nnz(void *h) { return (int64_t) static_cast<Result *>(h)->val.size(); }
int om_result_path(void *h) { return static_cast<Result *>(h)->path; }
void om_result_ops(void *h, double *out5) {
    const OpCount &o = static_cast<Result *>(h)->ops;
    out5[0] = o.flops;
    out5[1] = static_cast<double>(o.pairs);
    out5[2] = static_cast<double>(o.edges);
    out5[3] = static_cast<double>(o.intersections);
    out5[4] = static_cast<double>(o.fan_terms);
}
void om_result_stats(void *h, int64_t *out3) {
    Result *r = static_cast<Result *>(h);
    out3[0] = r->n_candidates;
    out3[1] = r->n_cap_rejected;
    out3[2] = r->n_clipped;
}
void om_result_csr(void *h, int64_t *row_ptr, int64_t *col, double *val) {
    Result *r = static_cast<Result *>(h);
    std::copy(r->row_ptr.begin(), r->row_ptr.end(), row_ptr);
    std::copy(r->col.begin(), r->col.end(), col);
    std::copy(r->val.begin(), r->val.end(), val);
}
void om_result_fracs(void *h, double *frac_a, double *frac_b, double *area_a, double *area_b) {
    Result *r = static_cast<Result *>(h);
    std::copy(r->frac_a.begin(), r->frac_a.end(), frac_a);
    std::copy(r->frac_b.begin(), r->frac_b.end(), frac_b);
    std::copy(r->area_a.begin(), r->area_a.end(), area_a);
    std::copy(r->area_b.begin(), r->area_b.end(), area_b);
}
int64_t om_result_n_candidates(void *h) { return (int64_t) static_cast<Result *>(h)->cand_idx.size(); }
void om_result_candidates(void *h, int64_t *offsets, int64_t *idx) {
    Result *r = static_cast<Result *>(h);
    std::copy(r->cand_offsets.begin(), r->cand_offsets.end(), offsets);
    std::copy(r->cand_idx.begin(), r->cand_idx.end(), idx);
}
void om_result_free(void *h) { delete static_cast<Result *>(h); }

// COO -> CSR with the reference's (row, col) sort.
void om_to_csr(const int64_t *rows, const int64_t *cols, const double *vals, int64_t nnz, int64_t n_rows, int64_t *row_ptr, int64_t *col,
               double *val) {
    Result r;
    r.n_dst = n_rows;
    std::vector<i64> R(rows, rows + nnz), C(cols, cols + nnz);
    std::vector<double> V(vals, vals + nnz);
    to_csr(r, R, C, V);
    std::copy(r.row_ptr.begin(), r.row_ptr.end(), row_ptr);
    std::copy(r.col.begin(), r.col.end(), col);
    std::copy(r.val.begin(), r.val.end(), val);
}

// ------------------------------------------------------------- apply
// CSR SpMV/SpMM, FP64, one row at a time with the nonzeros summed in CSR
// order (the KokkosKernels native kernel's per-row dot; apply.hpp:97, :306).
// Fields are layout_left [n, k].  n_threads threads split the rows.
void om_batch_apply_csr(const int64_t *row_ptr, const int64_t *col, const double *val, int64_t n_rows, int64_t n_src, const double *x,
                        double *y, int64_t k, int n_threads) {
    if (n_threads > 0) omp_set_num_threads(n_threads);
#pragma omp parallel for schedule(static)
    for (i64 r = 0; r < n_rows; ++r) {
        for (i64 f = 0; f < k; ++f) {
            double s = 0.0;
            for (i64 e = row_ptr[r]; e < row_ptr[r + 1]; ++e) s += val[e] * x[col[e] + f * n_src];
            y[r + f * n_rows] = s;
        }
    }
}

// nan_batch_apply, apply.hpp:359-420: per row the total weight Σ w (CSR
// order); per field the weighted sum and the valid-weight sum over non-NaN
// sources (CSR order); NaN when no valid weight, a zero total, or a valid
// fraction below 1 - na_thres - 1e-6; else weighted sum / valid weight.
void om_nan_batch_apply(const int64_t *row_ptr, const int64_t *col, const double *val, int64_t n_rows, int64_t n_src, const double *x,
                        double *y, int64_t k, double na_thres) {
    const double nan = std::numeric_limits<double>::quiet_NaN();
#pragma omp parallel for schedule(static)
    for (i64 r = 0; r < n_rows; ++r) {
        double total = 0.0;
        for (i64 e = row_ptr[r]; e < row_ptr[r + 1]; ++e) total += val[e];
        for (i64 f = 0; f < k; ++f) {
            double ws = 0.0, sv = 0.0;
            for (i64 e = row_ptr[r]; e < row_ptr[r + 1]; ++e) {
                const double v = x[col[e] + f * n_src];
                if (!std::isnan(v)) {
                    ws += val[e] * v;
                    sv += val[e];
                }
            }
            double out;
            if (sv == 0.0 || total == 0.0) {
                out = nan;
            } else {
                const double frac = sv / total;
                out = frac < (1.0 - na_thres - 1e-6) ? nan : ws / sv;
            }
            y[r + f * n_rows] = out;
        }
    }
}

// masked apply, apply.hpp:438-484: rows whose mask is 0 keep their dst
// value; the mask applies only when it has n_rows entries.  Sums in CSR order.
void om_masked_apply(const int64_t *row_ptr, const int64_t *col, const double *val, int64_t n_rows, int64_t n_src, const double *x, double *y,
                     int64_t k, const int32_t *mask, int64_t n_mask) {
    const bool has_mask = mask != nullptr && n_mask == n_rows;
#pragma omp parallel for schedule(static)
    for (i64 r = 0; r < n_rows; ++r) {
        if (has_mask && mask[r] == 0) continue;
        for (i64 f = 0; f < k; ++f) {
            double s = 0.0;
            for (i64 e = row_ptr[r]; e < row_ptr[r + 1]; ++e) s += val[e] * x[col[e] + f * n_src];
            y[r + f * n_rows] = s;
        }
    }
}

// check_conservation, conservation.cpp:82-103 (sequential sums).
void om_check_conservation(const double *src, int64_t n_src, const double *dst, int64_t n_dst, const double *area_a, const double *frac_a,
                           const double *area_b, const double *frac_b, int norm_type, double *out4) {
    double si = 0.0, di = 0.0;
    for (i64 i = 0; i < n_src; ++i) si += src[i] * area_a[i] * frac_a[i];
    if (norm_type == 0) {
        for (i64 j = 0; j < n_dst; ++j)
            if (frac_b[j] != 0.0) di += dst[j] * area_b[j];
    } else {
        for (i64 j = 0; j < n_dst; ++j) di += dst[j] * area_b[j] * frac_b[j];
    }
    double ab = std::abs(si - di);
    double den = std::max({std::abs(si), std::abs(di), std::numeric_limits<double>::epsilon()});
    out4[0] = si;
    out4[1] = di;
    out4[2] = ab;
    out4[3] = ab / den;
}

int om_max_threads() { return omp_get_max_threads(); }

}  // extern "C"
