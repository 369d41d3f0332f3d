/*
 * oa_bvh.h : host-side BVH builder (binned SAH, 16 bins) producing the
 * wide-2 BvhNode layout of oa_layout.h. Used for the TLAS over instance world
 * boxes and for one BLAS per solid over its CSGPrim boxes. This replaces the
 * OptiX GAS/IAS builds of CSGOptiX/SBT.cc:277-375,495-553.
 */
#pragma once
#include <stdint.h>
#include <vector>
#include <algorithm>
#include <cfloat>
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

    /* returns a child link for range [b,e) and its box */
    uint32_t link(std::vector<Item>& v, int b, int e, uint32_t depth, double* lo, double* hi)
    {
        box_of(v, b, e, lo, hi);
        if (e - b <= leaf_size) return leaf_link(v, b, e);
        return node(v, b, e, depth);
    }

    uint32_t node(std::vector<Item>& v, int b, int e, uint32_t depth)
    {
        uint32_t idx = (uint32_t)nodes.size();
        nodes.push_back(BvhNode());
        max_depth = std::max(max_depth, depth);
        int mid = split(v, b, e);
        double lo0[3], hi0[3], lo1[3], hi1[3];
        uint32_t c0 = link(v, b, mid, depth + 1, lo0, hi0);
        uint32_t c1 = link(v, mid, e, depth + 1, lo1, hi1);
        BvhNode& n = nodes[idx];
        store(n.lo0, n.hi0, lo0, hi0);
        store(n.lo1, n.hi1, lo1, hi1);
        n.c0 = c0; n.c1 = c1; n.depth = depth; n.pad = 0;
        return idx;
    }

    /* build a tree; returns the root node index (always an internal node record) */
    uint32_t build(std::vector<Item> v)
    {
        const int n = (int)v.size();
        if (n <= leaf_size) {
            uint32_t idx = (uint32_t)nodes.size();
            nodes.push_back(BvhNode());
            double lo[3], hi[3];
            box_of(v, 0, n, lo, hi);
            uint32_t c0 = leaf_link(v, 0, n);
            BvhNode& r = nodes[idx];
            store(r.lo0, r.hi0, lo, hi);
            for (int k = 0; k < 3; ++k) { r.lo1[k] = 1.f; r.hi1[k] = -1.f; }
            r.c0 = c0; r.c1 = OA_BVH_EMPTY; r.depth = 0; r.pad = 0;
            return idx;
        }
        return node(v, 0, n, 0);
    }
};

} // namespace oa_bvh
