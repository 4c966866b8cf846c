// geometry.hpp -- FP64 unit-sphere geometry, compiled for both the host and
// gfx950 (__host__ __device__).  Every expression keeps the operand order of
// the reference's SphericalClipper / SphericalPolygon
// (libs/axis/include/axis/detail/spherical_clipper.hpp:29-336) and the whole
// library is built with -ffp-contract=off, so + - * / sqrt give identical bits
// on the CPU and the GPU.  Only atan2 (polygon area) and acos (cap filter)
// differ by a few ulps between host libm and the device math library; the
// callers guard every decision that could flip on those ulps.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>

namespace axis_hip::geo {

struct V3 {
    double x, y, z;
};

__host__ __device__ inline double dot(const V3 &a, const V3 &b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__host__ __device__ inline V3 cross(const V3 &a, const V3 &b) {
    return V3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__host__ __device__ inline double length_sq(const V3 &v) { return v.x * v.x + v.y * v.y + v.z * v.z; }
__host__ __device__ inline V3 add(const V3 &a, const V3 &b) { return V3{a.x + b.x, a.y + b.y, a.z + b.z}; }
__host__ __device__ inline V3 normalize(const V3 &v) {
    double len = sqrt(length_sq(v));
    if (len < 1.0e-300) return V3{0.0, 0.0, 0.0};
    double inv = 1.0 / len;
    return V3{v.x * inv, v.y * inv, v.z * inv};
}

/// One Van Oosterom-Strackee fan term 2*atan2(a.(b x c), 1 + a.b + a.c + b.c)
/// (spherical_clipper.hpp:120-133).  `num_out` receives the numerator so a
/// caller can tell an exact zero from a rounding-level one.
__host__ __device__ inline double fan_term(const V3 &a, const V3 &b, const V3 &c, double &num_out) {
    double ab = dot(a, b);
    double ac = dot(a, c);
    double bc = dot(b, c);
    V3 bxc = cross(b, c);
    double num = dot(a, bxc);
    double den = 1.0 + ab + ac + bc;
    num_out = num;
    return 2.0 * atan2(num, den);
}

/// Intersection of arc v1->v2 with the great circle through (ea, eb),
/// spherical_clipper.hpp:263-286.
/// `n1` is the clip edge's normal cross(ea, eb), computed once per edge by the
/// callers (the same operands, so the same bits as computing it here).
__host__ __device__ inline V3 gc_intersection_n(const V3 &v1, const V3 &v2, const V3 &n1) {
    V3 n2 = cross(v1, v2);
    V3 d = cross(n1, n2);
    if (length_sq(d) < 1.0e-30) return normalize(add(v1, v2));
    V3 cand = normalize(d);
    V3 mid_dir = normalize(add(v1, v2));
    if (dot(cand, mid_dir) < 0.0) cand = V3{-cand.x, -cand.y, -cand.z};
    return cand;
}
__host__ __device__ inline V3 gc_intersection(const V3 &v1, const V3 &v2, const V3 &ea, const V3 &eb) {
    V3 n1 = cross(ea, eb);
    V3 n2 = cross(v1, v2);
    V3 d = cross(n1, n2);
    if (length_sq(d) < 1.0e-30) return normalize(add(v1, v2));
    V3 cand = normalize(d);
    V3 mid_dir = normalize(add(v1, v2));
    if (dot(cand, mid_dir) < 0.0) cand = V3{-cand.x, -cand.y, -cand.z};
    return cand;
}

/// Polygon storage with a compile-time capacity; the element accessors let the
/// same clipping code run over registers/stack on the host and over lane-private
/// LDS columns on the GPU.
struct PolyRef {
    double *x, *y, *z;  // base pointers
    int stride;         // element stride (1 on the host, wave width in LDS)
    __host__ __device__ V3 get(int i) const { return V3{x[i * stride], y[i * stride], z[i * stride]}; }
    __host__ __device__ void set(int i, const V3 &v) const {
        x[i * stride] = v.x;
        y[i * stride] = v.y;
        z[i * stride] = v.z;
    }
};

/// Clip `a` (subject, na vertices) against every edge of the clip polygon
/// given by `clip_vertex(k)` (nc vertices), Sutherland-Hodgman with the
/// inclusive -1e-14 half-space test (spherical_clipper.hpp:174-222, 301-335).
/// `b` is scratch of the same capacity.  The result lives in `a`.
///
/// `cap` is the buffer capacity.  The reference's SphericalPolygon<32>
/// silently drops vertices beyond 32; with cap == 32 this function does the
/// same.  With a smaller cap, a push that would exceed it sets *overflow and
/// stops, so the caller can redo the pair with capacity 32.
///
/// Returns the vertex count of the result (0 when it has fewer than 3).
template <class ClipVertex>
__host__ __device__ inline int clip_polygon(PolyRef a, int na, PolyRef b, ClipVertex clip_vertex, int nc, int cap, bool *overflow) {
    if (na < 3 || nc < 3) return 0;
    for (int ci = 0; ci < nc; ++ci) {
        const V3 ea = clip_vertex(ci);
        const V3 eb = clip_vertex(ci + 1 == nc ? 0 : ci + 1);
        V3 normal = cross(ea, eb);
        if (length_sq(normal) < 1.0e-30) continue;
        int nb = 0;
        constexpr double eps = -1.0e-14;
        // the reference's modulo walk: prev of vertex 0 is vertex na-1
        V3 prev = a.get(na - 1);
        double d_prev = dot(normal, prev);
        for (int i = 0; i < na; ++i) {
            V3 cur = a.get(i);
            double d_cur = dot(normal, cur);
            bool cur_in = d_cur >= eps;
            bool prev_in = d_prev >= eps;
            if (cur_in) {
                if (!prev_in) {
                    if (nb < cap) {
                        b.set(nb++, gc_intersection(prev, cur, ea, eb));
                    } else if (cap < 32) {
                        *overflow = true;
                        return 0;
                    }
                }
                if (nb < cap) {
                    b.set(nb++, cur);
                } else if (cap < 32) {
                    *overflow = true;
                    return 0;
                }
            } else if (prev_in) {
                if (nb < cap) {
                    b.set(nb++, gc_intersection(prev, cur, ea, eb));
                } else if (cap < 32) {
                    *overflow = true;
                    return 0;
                }
            }
            prev = cur;
            d_prev = d_cur;
        }
        for (int i = 0; i < nb; ++i) a.set(i, b.get(i));
        na = nb;
        if (na < 3) return 0;
    }
    return na;
}

/// The inclusive half-space test of the clipper (spherical_clipper.hpp:
/// 307-319): a vertex at signed distance d from the clip plane is kept when
/// d >= -1e-14.
__host__ __device__ inline bool inside_half(double d) { return d >= -1.0e-14; }

/// One clip edge of clip_polygon on one buffer: the walk over the subject's
/// vertices reads one vertex ahead, and each step writes its (at most two)
/// results at positions the walk has already read, which holds whenever the
/// subject enters the clip half-space at most once per edge -- always for the
/// convex cells a valid mesh has.  Anything else (a write that would land
/// ahead of the walk, or more than `cap` vertices) sets *overflow.  `normal`
/// is the clip edge's cross(ea, eb).  Returns the new vertex count.
__host__ __device__ inline int clip_edge_walk(PolyRef a, int na, const V3 &normal, int cap, bool *overflow) {
    int nb = 0;
    V3 prev = a.get(na - 1);
    double d_prev = dot(normal, prev);
    V3 next = a.get(0);
    for (int i = 0; i < na; ++i) {
        const V3 cur = next;
        if (i + 1 < na) next = a.get(i + 1);
        const double d_cur = dot(normal, cur);
        const bool cur_in = inside_half(d_cur);
        const bool prev_in = inside_half(d_prev);
        const int emit = cur_in ? (prev_in ? 1 : 2) : (prev_in ? 1 : 0);
        if (emit > 0 && (nb + emit - 1 > i + 1 || nb + emit > cap)) {
            *overflow = true;
            return 0;
        }
        if (cur_in) {
            if (!prev_in) a.set(nb++, gc_intersection_n(prev, cur, normal));
            a.set(nb++, cur);
        } else if (prev_in) {
            a.set(nb++, gc_intersection_n(prev, cur, normal));
        }
        prev = cur;
        d_prev = d_cur;
    }
    return nb;
}

/// clip_polygon on one buffer (clip_edge_walk per clip edge): on *overflow
/// the caller decides the pair with clip_polygon on the host.  The
/// arithmetic and the order of the results are those of clip_polygon.
template <class ClipVertex>
__host__ __device__ inline int clip_polygon_inplace(PolyRef a, int na, ClipVertex clip_vertex, int nc, int cap, bool *overflow) {
    if (na < 3 || nc < 3) return 0;
    for (int ci = 0; ci < nc; ++ci) {
        const V3 ea = clip_vertex(ci);
        const V3 eb = clip_vertex(ci + 1 == nc ? 0 : ci + 1);
        V3 normal = cross(ea, eb);
        if (length_sq(normal) < 1.0e-30) continue;
        na = clip_edge_walk(a, na, normal, cap, overflow);
        if (*overflow || na < 3) return 0;
    }
    return na;
}

/// clip_polygon_inplace arranged for lockstep lanes.  Per clip edge: the
/// side of every vertex first (a bit mask), then the intersections at the
/// two places where a convex subject crosses the edge, then the surviving
/// vertices moved into place.  A wave then runs the intersection code twice
/// per edge, where the vertex walk runs it at every position at which any of
/// its lanes crosses.  An edge that the subject crosses more than twice (not
/// convex) takes clip_edge_walk.  Every intersection has the walk's operands
/// and the output its order, so the results are clip_polygon_inplace's bit
/// for bit, overflow included.  Needs cap < 32.
template <class ClipVertex>
__host__ __device__ inline int clip_polygon_slots(PolyRef a, int na, ClipVertex clip_vertex, int nc, int cap, bool *overflow) {
    if (na < 3 || nc < 3) return 0;
    if (na > cap) {
        *overflow = true;
        return 0;
    }
    for (int ci = 0; ci < nc; ++ci) {
        const V3 ea = clip_vertex(ci);
        const V3 eb = clip_vertex(ci + 1 == nc ? 0 : ci + 1);
        const V3 normal = cross(ea, eb);
        if (length_sq(normal) < 1.0e-30) continue;
        unsigned in = 0;  // bit i: vertex i inside (na <= cap < 32)
        for (int i = 0; i < na; ++i) in |= static_cast<unsigned>(inside_half(dot(normal, a.get(i)))) << i;
        if (in == 0) return 0;  // every vertex outside: nothing is emitted
        const unsigned full = (1u << na) - 1u;
        const unsigned prev_in = ((in << 1) | (in >> (na - 1))) & full;  // bit i: vertex i-1 (cyclic) inside
        const unsigned x = in ^ prev_in;                                // bit i: the arc (i-1, i) crosses
        if (x == 0) continue;  // every vertex inside: the walk rewrites the polygon unchanged
        if (__builtin_popcount(x) != 2) {
            na = clip_edge_walk(a, na, normal, cap, overflow);
            if (*overflow || na < 3) return 0;
            continue;
        }
        const int i0 = __builtin_ctz(x), i1 = __builtin_ctz(x & (x - 1u));
        const int nb = __builtin_popcount(in) + 2;
        if (nb > cap) {
            *overflow = true;
            return 0;
        }
        const V3 x0 = gc_intersection_n(a.get(i0 == 0 ? na - 1 : i0 - 1), a.get(i0), normal);
        const V3 x1 = gc_intersection_n(a.get(i1 - 1), a.get(i1), normal);
        if ((in >> i0) & 1u) {
            // enters at i0, leaves at i1: X0, v[i0 .. i1-1], X1
            const int m = i1 - i0;
            if (i0 == 0) {
                for (int k = m - 1; k >= 0; --k) a.set(k + 1, a.get(k));
            } else {
                for (int k = 0; k < m; ++k) a.set(k + 1, a.get(i0 + k));
            }
            a.set(0, x0);
            a.set(m + 1, x1);
        } else {
            // leaves at i0, enters at i1: v[0 .. i0-1], X0, X1, v[i1 .. na-1]
            const int t = na - i1;
            if (i1 == i0 + 1) {
                for (int k = t - 1; k >= 0; --k) a.set(i0 + 2 + k, a.get(i1 + k));
            } else {
                for (int k = 0; k < t; ++k) a.set(i0 + 2 + k, a.get(i1 + k));
            }
            a.set(i0, x0);
            a.set(i0 + 1, x1);
        }
        na = nb;
    }
    return na;
}

/// SphericalPolygon::area (spherical_clipper.hpp:112-137).  Also returns the
/// sum of |terms| and whether every numerator was exactly zero, so a caller
/// can detect cancellation that a different atan2 could resolve differently.
__host__ __device__ inline double polygon_area(PolyRef p, int n, double *abs_sum, bool *all_zero_num) {
    *abs_sum = 0.0;
    *all_zero_num = true;
    if (n < 3) return 0.0;
    double total = 0.0;
    V3 a = p.get(0);
    V3 b = p.get(1);
    for (int i = 1; i < n - 1; ++i) {
        V3 c = p.get(i + 1);
        double num;
        double t = fan_term(a, b, c, num);
        if (num != 0.0) *all_zero_num = false;
        total += t;
        *abs_sum += fabs(t);
        b = c;
    }
    return fabs(total);
}

}  // namespace axis_hip::geo
