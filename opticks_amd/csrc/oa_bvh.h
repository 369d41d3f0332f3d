/*
 * oa_bvh.h : host-side BVH builder producing the 4-wide BvhNode layout of
 * oa_layout.h. Binary binned-SAH splits (16 bins, plus the "largest item
 * alone" candidate) are collapsed four-wide: a node starts from the two halves
 * of its range and splits the child of largest surface area until it has four
 * children. Used for the TLAS over instance world boxes and for one BLAS per
 * solid over its CSGPrim boxes. This replaces the OptiX GAS/IAS builds of
 * CSGOptiX/SBT.cc:277-375,495-553.
 *
 * stack_need: the deepest traversal stack a nearest-first walk can reach below
 * the root, sum over a root-to-leaf path of (children - 1), each node pushing
 * all hit children but the nearest.
 */
#pragma once
#include <stdint.h>
#include <vector>
#include <algorithm>
#include <cfloat>
#include <cmath>
#include "oa_layout.h"

namespace oa_bvh {

struct Item {
    double lo[3], hi[3];
    uint32_t id;
    double c(int k) const { return 0.5 * (lo[k] + hi[k]); }
};

struct Builder {
    std::vector<BvhNode>& nodes;
    std::vector<uint32_t>& items;
    int leaf_size;
    uint32_t max_depth = 0;
    uint32_t stack_need = 0;

    Builder(std::vector<BvhNode>& n, std::vector<uint32_t>& it, int leaf) : nodes(n), items(it), leaf_size(leaf) {}

    static void box_of(const std::vector<Item>& v, int b, int e, double* lo, double* hi)
    {
        for (int k = 0; k < 3; ++k) { lo[k] = DBL_MAX; hi[k] = -DBL_MAX; }
        for (int i = b; i < e; ++i)
            for (int k = 0; k < 3; ++k) { lo[k] = std::min(lo[k], v[i].lo[k]); hi[k] = std::max(hi[k], v[i].hi[k]); }
    }

    static double area(const double* lo, const double* hi)
    {
        double d0 = std::max(0.0, hi[0] - lo[0]), d1 = std::max(0.0, hi[1] - lo[1]), d2 = std::max(0.0, hi[2] - lo[2]);
        return d0 * d1 + d1 * d2 + d2 * d0;
    }

    /* outward rounding so float boxes still contain the double boxes */
    static void store(float* flo, float* fhi, const double* lo, const double* hi)
    {
        for (int k = 0; k < 3; ++k) {
            float a = (float)lo[k], b = (float)hi[k];
            if ((double)a > lo[k]) a = nextafterf(a, -FLT_MAX);
            if ((double)b < hi[k]) b = nextafterf(b, FLT_MAX);
            flo[k] = a; fhi[k] = b;
        }
    }

    uint32_t leaf_link(std::vector<Item>& v, int b, int e)
    {
        uint32_t first = (uint32_t)items.size();
        for (int i = b; i < e; ++i) items.push_back(v[i].id);
        return OA_BVH_LEAF_BIT | ((uint32_t)(e - b) << 24) | (first & 0x00ffffffu);
    }

    int split(std::vector<Item>& v, int b, int e)
    {
        double clo[3] = { DBL_MAX, DBL_MAX, DBL_MAX }, chi[3] = { -DBL_MAX, -DBL_MAX, -DBL_MAX };
        for (int i = b; i < e; ++i)
            for (int k = 0; k < 3; ++k) { clo[k] = std::min(clo[k], v[i].c(k)); chi[k] = std::max(chi[k], v[i].c(k)); }
        const int NB = 16;
        double best_cost = DBL_MAX; int best_axis = -1, best_bin = -1;
        for (int ax = 0; ax < 3; ++ax) {
            double ext = chi[ax] - clo[ax];
            if (ext <= 0.0) continue;
            int cnt[NB] = {};
            double blo[NB][3], bhi[NB][3];
            for (int j = 0; j < NB; ++j) for (int k = 0; k < 3; ++k) { blo[j][k] = DBL_MAX; bhi[j][k] = -DBL_MAX; }
            for (int i = b; i < e; ++i) {
                int j = std::min(NB - 1, (int)(NB * (v[i].c(ax) - clo[ax]) / ext));
                cnt[j]++;
                for (int k = 0; k < 3; ++k) { blo[j][k] = std::min(blo[j][k], v[i].lo[k]); bhi[j][k] = std::max(bhi[j][k], v[i].hi[k]); }
            }
            for (int s = 1; s < NB; ++s) {
                double llo[3] = { DBL_MAX, DBL_MAX, DBL_MAX }, lhi[3] = { -DBL_MAX, -DBL_MAX, -DBL_MAX };
                double rlo[3] = { DBL_MAX, DBL_MAX, DBL_MAX }, rhi[3] = { -DBL_MAX, -DBL_MAX, -DBL_MAX };
                int nl = 0, nr = 0;
                for (int j = 0; j < NB; ++j) {
                    double* lo = j < s ? llo : rlo; double* hi = j < s ? lhi : rhi;
                    if (!cnt[j]) continue;
                    (j < s ? nl : nr) += cnt[j];
                    for (int k = 0; k < 3; ++k) { lo[k] = std::min(lo[k], blo[j][k]); hi[k] = std::max(hi[k], bhi[j][k]); }
                }
                if (!nl || !nr) continue;
                double cost = nl * area(llo, lhi) + nr * area(rlo, rhi);
                if (cost < best_cost) { best_cost = cost; best_axis = ax; best_bin = s; }
            }
        }
        /* candidate: the largest-area item alone (an enclosing world box binned
           by its centroid would otherwise inflate one subtree to the whole scene) */
        int big = b;
        double big_area = -1.0;
        for (int i = b; i < e; ++i) {
            double a = area(v[i].lo, v[i].hi);
            if (a > big_area) { big_area = a; big = i; }
        }
        {
            double rlo[3] = { DBL_MAX, DBL_MAX, DBL_MAX }, rhi[3] = { -DBL_MAX, -DBL_MAX, -DBL_MAX };
            for (int i = b; i < e; ++i) {
                if (i == big) continue;
                for (int k = 0; k < 3; ++k) { rlo[k] = std::min(rlo[k], v[i].lo[k]); rhi[k] = std::max(rhi[k], v[i].hi[k]); }
            }
            double cost = big_area + (e - b - 1) * area(rlo, rhi);
            if (cost < best_cost) {
                std::swap(v[b], v[big]);
                return b + 1;
            }
        }
        int mid;
        if (best_axis < 0) {
            mid = b + (e - b) / 2;
        } else {
            const double ext = chi[best_axis] - clo[best_axis];
            auto it = std::partition(v.begin() + b, v.begin() + e, [&](const Item& x) {
                int j = std::min(NB - 1, (int)(NB * (x.c(best_axis) - clo[best_axis]) / ext));
                return j < best_bin;
            });
            mid = (int)(it - v.begin());
            if (mid == b || mid == e) mid = b + (e - b) / 2;
        }
        return mid;
    }

    /* returns a child link for range [b,e), its box and the stack need below it */
    uint32_t link(std::vector<Item>& v, int b, int e, uint32_t depth, double* lo, double* hi, uint32_t& need)
    {
        box_of(v, b, e, lo, hi);
        need = 0;
        if (e - b <= leaf_size) return leaf_link(v, b, e);
        return node(v, b, e, depth, need);
    }

    uint32_t node(std::vector<Item>& v, int b, int e, uint32_t depth, uint32_t& need)
    {
        uint32_t idx = (uint32_t)nodes.size();
        nodes.push_back(BvhNode());
        max_depth = std::max(max_depth, depth);
        /* ranges of the children: split the largest splittable one until four */
        int rb[OA_BVH_WIDTH], re[OA_BVH_WIDTH], nc = 1;
        rb[0] = b; re[0] = e;
        while (nc < OA_BVH_WIDTH) {
            int pick = -1;
            double best = -1.0;
            for (int c = 0; c < nc; ++c) {
                if (re[c] - rb[c] <= leaf_size) continue;
                double lo[3], hi[3];
                box_of(v, rb[c], re[c], lo, hi);
                const double a = area(lo, hi);
                if (a > best) { best = a; pick = c; }
            }
            if (pick < 0) break;
            const int mid = split(v, rb[pick], re[pick]);
            rb[nc] = mid; re[nc] = re[pick];
            re[pick] = mid;
            ++nc;
        }
        /* leaves first: on equal entry distance the traversal keeps this order,
           so a large enclosing leaf (a world volume) is tested before subtrees */
        int order[OA_BVH_WIDTH];
        for (int c = 0; c < nc; ++c) order[c] = c;
        std::stable_sort(order, order + nc, [&](int x, int y) {
            return (re[x] - rb[x] <= leaf_size) > (re[y] - rb[y] <= leaf_size);
        });
        uint32_t links[OA_BVH_WIDTH];
        double clo[OA_BVH_WIDTH][3], chi[OA_BVH_WIDTH][3];
        uint32_t below = 0;
        for (int k = 0; k < nc; ++k) {
            uint32_t n_below = 0;
            const int c = order[k];
            links[k] = link(v, rb[c], re[c], depth + 1, clo[k], chi[k], n_below);
            below = std::max(below, n_below);
        }
        need = (uint32_t)(nc - 1) + below;
        BvhNode& n = nodes[idx];
        fill(n, nc, links, clo, chi);
        n.depth = depth;
        return idx;
    }

    static void fill(BvhNode& n, int nc, const uint32_t* links, const double (*lo)[3], const double (*hi)[3])
    {
        for (int k = 0; k < OA_BVH_WIDTH; ++k) {
            if (k < nc) {
                float flo[3], fhi[3];
                store(flo, fhi, lo[k], hi[k]);
                n.lox[k] = flo[0]; n.loy[k] = flo[1]; n.loz[k] = flo[2];
                n.hix[k] = fhi[0]; n.hiy[k] = fhi[1]; n.hiz[k] = fhi[2];
                n.child[k] = links[k];
            } else {                    /* empty slot: a box at +inf, which every slab test misses */
                n.lox[k] = n.loy[k] = n.loz[k] = INFINITY;
                n.hix[k] = n.hiy[k] = n.hiz[k] = INFINITY;
                n.child[k] = OA_BVH_EMPTY;
            }
        }
        n.pad[0] = n.pad[1] = n.pad[2] = 0;
    }

    /* build a tree; returns the root node index (always an internal node record) */
    uint32_t build(std::vector<Item> v)
    {
        const int n = (int)v.size();
        if (n == 0) {                   /* no items: a root whose four slots are empty */
            uint32_t idx = (uint32_t)nodes.size();
            nodes.push_back(BvhNode());
            fill(nodes[idx], 0, nullptr, nullptr, nullptr);
            nodes[idx].depth = 0;
            return idx;
        }
        if (n <= leaf_size) {
            uint32_t idx = (uint32_t)nodes.size();
            nodes.push_back(BvhNode());
            double lo[1][3], hi[1][3];
            box_of(v, 0, n, lo[0], hi[0]);
            uint32_t c0 = leaf_link(v, 0, n);
            fill(nodes[idx], 1, &c0, lo, hi);
            nodes[idx].depth = 0;
            return idx;
        }
        uint32_t need = 0;
        const uint32_t root = node(v, 0, n, 0, need);
        stack_need = std::max(stack_need, need);
        return root;
    }
};

} // namespace oa_bvh
