/*
 * oa_bounds.h : host-side world bounds of CSG prims under an instance transform,
 * for the TLAS leaf boxes.
 *
 * The box of a rotated prim built from the corners of its solid-frame AABB can
 * be several times larger than the prim itself (a detector tilted on a
 * spherical shell). Here each leaf is bounded exactly through its support
 * function in world space, with the node transform and the instance transform
 * composed:
 *   sphere (c, r)           centre A c + t, half extent r |row_i(A)|
 *   zsphere                 the sphere's box intersected with the box of its z slab
 *   cylinder (r, z1..z2)    disk (r, in the xy plane) plus the axis segment:
 *                           t_i +- r |(A_i0, A_i1)| + (min|max)(z1 A_i2, z2 A_i2)
 *   box3 (full sides)       its eight corners
 * and combined over the tree: union -> union of boxes, intersection ->
 * intersection, difference -> the left operand's box. Complemented nodes, list
 * nodes and the other leaf types return "unbounded"; the caller intersects the
 * result with the transformed prim AABB, so they fall back to that box. Boxes
 * only prune traversal: any conservative box gives the same hits.
 *
 * Also the front-instance split and the free ball of the TLAS (below).
 *
 * Conventions: CSGNode q0..q3 as the device reads them (sphere q0 = centre,
 * radius; zsphere q1.x/q1.y = z cuts relative to the centre; cylinder q0.w =
 * radius, q1.x/q1.y = z range; box3 q0.xyz = full sides; q3.z typecode, q3.w
 * transform index + 1 with the complement in bit 31). itra rows map the prim
 * frame to the leaf frame as the device's q_right does (row vector times the
 * 4x4); instance transforms (o2w) map object to world the same way.
 */
#pragma once
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <utility>
#include <vector>

namespace oa_bounds {

struct Box {
    double lo[3], hi[3];
};

inline Box unbounded()
{
    Box b;
    for (int k = 0; k < 3; ++k) { b.lo[k] = -1e300; b.hi[k] = 1e300; }
    return b;
}
inline Box empty()
{
    Box b;
    for (int k = 0; k < 3; ++k) { b.lo[k] = 1e300; b.hi[k] = -1e300; }
    return b;
}
inline Box join(const Box& a, const Box& b)
{
    Box r;
    for (int k = 0; k < 3; ++k) { r.lo[k] = std::min(a.lo[k], b.lo[k]); r.hi[k] = std::max(a.hi[k], b.hi[k]); }
    return r;
}
inline Box meet(const Box& a, const Box& b)
{
    Box r;
    for (int k = 0; k < 3; ++k) { r.lo[k] = std::max(a.lo[k], b.lo[k]); r.hi[k] = std::min(a.hi[k], b.hi[k]); }
    return r;
}

/* column-form affine map p' = L p + t */
struct Affine {
    double L[3][3], t[3];
};

/* from a row-vector 4x4 (p'_k = sum_j p_j m[4j+k] + m[12+k]) */
inline Affine from_rows(const float* m)
{
    Affine a;
    for (int k = 0; k < 3; ++k) {
        for (int j = 0; j < 3; ++j) a.L[k][j] = (double)m[4 * j + k];
        a.t[k] = (double)m[12 + k];
    }
    return a;
}

inline Affine compose(const Affine& outer, const Affine& inner)
{
    Affine r;
    for (int i = 0; i < 3; ++i) {
        for (int j = 0; j < 3; ++j) {
            double s = 0.0;
            for (int k = 0; k < 3; ++k) s += outer.L[i][k] * inner.L[k][j];
            r.L[i][j] = s;
        }
        double s = outer.t[i];
        for (int k = 0; k < 3; ++k) s += outer.L[i][k] * inner.t[k];
        r.t[i] = s;
    }
    return r;
}

/* inverse of an affine map; false when singular */
inline bool invert(const Affine& a, Affine& r)
{
    const double (*m)[3] = a.L;
    const double c00 = m[1][1] * m[2][2] - m[1][2] * m[2][1];
    const double c01 = m[1][2] * m[2][0] - m[1][0] * m[2][2];
    const double c02 = m[1][0] * m[2][1] - m[1][1] * m[2][0];
    const double det = m[0][0] * c00 + m[0][1] * c01 + m[0][2] * c02;
    if (!(std::fabs(det) > 0.0) || !std::isfinite(det)) return false;
    r.L[0][0] = c00 / det;
    r.L[0][1] = (m[0][2] * m[2][1] - m[0][1] * m[2][2]) / det;
    r.L[0][2] = (m[0][1] * m[1][2] - m[0][2] * m[1][1]) / det;
    r.L[1][0] = c01 / det;
    r.L[1][1] = (m[0][0] * m[2][2] - m[0][2] * m[2][0]) / det;
    r.L[1][2] = (m[0][2] * m[1][0] - m[0][0] * m[1][2]) / det;
    r.L[2][0] = c02 / det;
    r.L[2][1] = (m[0][1] * m[2][0] - m[0][0] * m[2][1]) / det;
    r.L[2][2] = (m[0][0] * m[1][1] - m[0][1] * m[1][0]) / det;
    for (int i = 0; i < 3; ++i) r.t[i] = -(r.L[i][0] * a.t[0] + r.L[i][1] * a.t[1] + r.L[i][2] * a.t[2]);
    return true;
}

/* world box of a local axis-aligned box under A */
inline Box box_of_box(const Affine& A, const double lo[3], const double hi[3])
{
    Box b;
    for (int i = 0; i < 3; ++i) {
        double c = A.t[i], h = 0.0;
        for (int j = 0; j < 3; ++j) {
            c += A.L[i][j] * 0.5 * (lo[j] + hi[j]);
            h += std::fabs(A.L[i][j]) * 0.5 * (hi[j] - lo[j]);
        }
        b.lo[i] = c - h;
        b.hi[i] = c + h;
    }
    return b;
}

inline uint32_t word(const float* f) { uint32_t u; std::memcpy(&u, f, 4); return u; }

/* world box of one leaf node (16 floats) under A (leaf frame -> world) */
inline Box leaf_box(const float* nd, const Affine& A)
{
    const uint32_t tc = word(nd + 14);
    if (tc == 101u || tc == 103u) {                          /* sphere, zsphere */
        const double c[3] = { nd[0], nd[1], nd[2] }, r = nd[3];
        Box s;
        for (int i = 0; i < 3; ++i) {
            double ci = A.t[i], n2 = 0.0;
            for (int j = 0; j < 3; ++j) { ci += A.L[i][j] * c[j]; n2 += A.L[i][j] * A.L[i][j]; }
            const double h = r * std::sqrt(n2);
            s.lo[i] = ci - h;
            s.hi[i] = ci + h;
        }
        if (tc == 103u) {
            const double lo[3] = { c[0] - r, c[1] - r, c[2] + std::max((double)nd[4], -r) };
            const double hi[3] = { c[0] + r, c[1] + r, c[2] + std::min((double)nd[5], r) };
            s = meet(s, box_of_box(A, lo, hi));
        }
        return s;
    }
    if (tc == 105u) {                                        /* cylinder */
        const double r = nd[3], z1 = nd[4], z2 = nd[5];
        Box b;
        for (int i = 0; i < 3; ++i) {
            const double rad = r * std::sqrt(A.L[i][0] * A.L[i][0] + A.L[i][1] * A.L[i][1]);
            const double a = z1 * A.L[i][2], e = z2 * A.L[i][2];
            b.lo[i] = A.t[i] - rad + std::min(a, e);
            b.hi[i] = A.t[i] + rad + std::max(a, e);
        }
        return b;
    }
    if (tc == 110u) {                                        /* box3 */
        const double lo[3] = { -0.5 * nd[0], -0.5 * nd[1], -0.5 * nd[2] };
        const double hi[3] = { 0.5 * nd[0], 0.5 * nd[1], 0.5 * nd[2] };
        return box_of_box(A, lo, hi);
    }
    return unbounded();
}

/* world box of node `idx` (1-based, complete-tree order) of the prim rooted at
   `root` (num_tree tree nodes), under M (prim frame -> world) */
inline Box node_box(const float* root, uint32_t idx, uint32_t num_tree, const float* itra, int64_t n_tran,
                    const Affine& M, int depth)
{
    if (idx == 0u || idx > num_tree || depth > 16) return unbounded();
    const float* nd = root + 16 * (idx - 1u);
    const uint32_t tc = word(nd + 14);
    const uint32_t gtw = word(nd + 15);
    if (tc == 0u) return empty();
    if (gtw & 0x80000000u) return unbounded();               /* complemented: not bounded */
    if (tc >= 1u && tc <= 3u) {
        const Box l = node_box(root, 2u * idx, num_tree, itra, n_tran, M, depth + 1);
        if (tc == 3u) return l;                              /* difference lies inside its left operand */
        const Box r = node_box(root, 2u * idx + 1u, num_tree, itra, n_tran, M, depth + 1);
        return tc == 1u ? join(l, r) : meet(l, r);
    }
    if (tc < 101u) return unbounded();                       /* list nodes */
    Affine A = M;
    const uint32_t gt = gtw & 0x7fffffffu;
    if (gt > 0u) {
        if (!itra || (int64_t)gt > n_tran) return unbounded();
        Affine inv, fwd;
        inv = from_rows(itra + 16 * (gt - 1u));              /* prim frame -> leaf frame */
        if (!invert(inv, fwd)) return unbounded();
        A = compose(M, fwd);
    }
    return leaf_box(nd, A);
}

/* world box of a prim (root CSGNode at `root`) under M */
inline Box prim_box(const float* root, const float* itra, int64_t n_tran, const Affine& M)
{
    const uint32_t tc = word(root + 14);
    const uint32_t num = (tc >= 1u && tc <= 3u) ? word(root) : 1u;
    return node_box(root, 1u, num, itra, n_tran, M, 0);
}

/* world box of an instance (o2w: 16 floats, row vector) of a solid whose prims are
   prim[first .. first+count) (CSGPrim records, 16 floats: node offset at word 1,
   AABB at 8..13): per prim the exact bound met with its transformed AABB, joined
   over the prims; unpadded */
inline Box instance_box(const float* o2w, const float* node, const float* itra, int64_t n_tran, const float* prim,
                        int32_t first, int32_t count)
{
    const Affine M = from_rows(o2w);
    Box all = empty();
    for (int32_t p = first; p < first + count; ++p) {
        const float* rec = prim + 16 * (int64_t)p;
        const double lo[3] = { rec[8], rec[9], rec[10] }, hi[3] = { rec[11], rec[12], rec[13] };
        int32_t off;
        std::memcpy(&off, rec + 1, 4);
        all = join(all, meet(box_of_box(M, lo, hi), prim_box(node + 16 * (int64_t)off, itra, n_tran, M)));
    }
    return all;
}

inline double area(const Box& b)
{
    const double x = b.hi[0] - b.lo[0], y = b.hi[1] - b.lo[1], z = b.hi[2] - b.lo[2];
    return x * y + y * z + z * x;
}

/* Front instances: the (at most max_front) instances whose world box has at least
   a quarter of the surface area of the box around all instances -- the world
   volume and the large shells inside it, which nearly every ray meets. */
inline std::vector<char> front_mask(const std::vector<Box>& boxes, size_t max_front)
{
    Box u = empty();
    for (const Box& b : boxes) u = join(u, b);
    const double ua = boxes.empty() ? 0.0 : area(u);
    std::vector<std::pair<double, size_t>> big;
    for (size_t i = 0; i < boxes.size(); ++i)
        if (area(boxes[i]) >= 0.25 * ua) big.push_back({ -area(boxes[i]), i });
    std::sort(big.begin(), big.end());
    if (big.size() > max_front) big.resize(max_front);
    std::vector<char> front(boxes.size(), 0);
    for (const auto& b : big) front[b.second] = 1;
    return front;
}

/* distance from point c to box b (0 inside) */
inline double box_distance(const double c[3], const Box& b)
{
    double d2 = 0.0;
    for (int k = 0; k < 3; ++k) {
        const double e = std::max(std::max(b.lo[k] - c[k], c[k] - b.hi[k]), 0.0);
        d2 += e * e;
    }
    return std::sqrt(d2);
}

/* distance from point c to the box `local` (instance frame) placed by o2w, when o2w is
   rigid (orthonormal linear part: distances are kept); 0 otherwise */
inline double placed_box_distance(const double c[3], const Box& local, const float* o2w)
{
    const Affine A = from_rows(o2w);
    double err = 0.0;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            double s = 0.0;
            for (int k = 0; k < 3; ++k) s += A.L[k][i] * A.L[k][j];
            err = std::max(err, std::fabs(s - (i == j ? 1.0 : 0.0)));
        }
    Affine inv;
    if (err > 1e-6 || !invert(A, inv)) return 0.0;
    double cl[3];
    for (int i = 0; i < 3; ++i) cl[i] = inv.t[i] + inv.L[i][0] * c[0] + inv.L[i][1] * c[1] + inv.L[i][2] * c[2];
    return box_distance(cl, local);
}

/* A ball that meets no instance not marked front, less a margin for float arithmetic
   on points up to |c| + r from the origin; r2 <= 0: none. The distance to an instance
   is the larger of the distances to its world box and (with o2w, 16 floats per
   instance, and locals, its box in its own frame) to its placed local box: both are
   lower bounds of the distance to its shape. The centre is the centre of the front
   boxes' union or of all boxes, whichever gives the larger radius. */
struct Ball {
    double c[3], r2;
};

inline Ball free_ball(const std::vector<Box>& boxes, const std::vector<char>& front,
                      const std::vector<Box>* locals = nullptr, const float* o2w = nullptr)
{
    Box fu = empty(), au = empty();
    bool any_front = false, any_rest = false;
    for (size_t i = 0; i < boxes.size(); ++i) {
        au = join(au, boxes[i]);
        if (front[i]) { fu = join(fu, boxes[i]); any_front = true; }
        else any_rest = true;
    }
    Ball best = { { 0.0, 0.0, 0.0 }, -1.0 };
    if (!any_rest) return best;
    double best_r = 0.0;
    for (int which = 0; which < 2; ++which) {
        if (which == 0 && !any_front) continue;
        const Box& u = which == 0 ? fu : au;
        double c[3];
        for (int k = 0; k < 3; ++k) c[k] = 0.5 * (u.lo[k] + u.hi[k]);
        double r = 1e300;
        for (size_t i = 0; i < boxes.size(); ++i) {
            if (front[i]) continue;
            double d = box_distance(c, boxes[i]);
            if (locals && o2w) d = std::max(d, placed_box_distance(c, (*locals)[i], o2w + 16 * i));
            r = std::min(r, d);
        }
        const double cn = std::sqrt(c[0] * c[0] + c[1] * c[1] + c[2] * c[2]);
        const double rm = r - 1e-4 * (r + cn) - 1e-2;
        if (rm > best_r) {
            best_r = rm;
            best = { { c[0], c[1], c[2] }, rm * rm };
        }
    }
    return best;
}

} // namespace oa_bounds
