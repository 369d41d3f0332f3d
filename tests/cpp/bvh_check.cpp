// Host check of the 4-wide BVH builder (opticks_amd/csrc/oa_bvh.h) on boxes read
// from a binary file of (n,6) doubles (lo.xyz, hi.xyz): every item in exactly one
// leaf, every child box inside its parent's slot, the stack bound, and a
// nearest-first walk with the kernel's slab form (fma with finite reciprocals)
// that reaches every box the segment [0, tmax] enters by more than 1e-3.
// The float4/uint4 records below stand in for HIP's vector types, which the
// builder does not use.
struct float4 { float x, y, z, w; };
struct uint4 { unsigned x, y, z, w; };
#include "oa_bvh.h"
#include <cmath>
#include <cstdio>
#include <random>
#include <set>
#include <vector>

int main(int argc, char** argv)
{
    if (argc < 4) return 2;
    FILE* f = std::fopen(argv[1], "rb");
    if (!f) return 2;
    std::vector<double> b;
    double x;
    while (std::fread(&x, 8, 1, f) == 1) b.push_back(x);
    std::fclose(f);
    const int n = (int)b.size() / 6, leaf = std::atoi(argv[2]), n_rays = std::atoi(argv[3]);
    std::vector<oa_bvh::Item> items(n);
    for (int i = 0; i < n; ++i) {
        for (int k = 0; k < 3; ++k) { items[i].lo[k] = b[6 * i + k]; items[i].hi[k] = b[6 * i + 3 + k]; }
        items[i].id = (uint32_t)i;
    }
    std::vector<BvhNode> nodes;
    std::vector<uint32_t> its;
    oa_bvh::Builder B(nodes, its, leaf);
    const uint32_t root = B.build(items);
    std::vector<int> seen(n, 0);
    long bad_box = 0;
    std::vector<uint32_t> todo{root};
    while (!todo.empty()) {
        const BvhNode& nd = nodes[todo.back()];
        todo.pop_back();
        for (int c = 0; c < OA_BVH_WIDTH; ++c) {
            const uint32_t l = nd.child[c];
            if (l == OA_BVH_EMPTY) continue;
            const float lo[3] = { nd.lox[c], nd.loy[c], nd.loz[c] }, hi[3] = { nd.hix[c], nd.hiy[c], nd.hiz[c] };
            if (l & OA_BVH_LEAF_BIT) {
                const uint32_t fi = l & 0xffffffu, cnt = (l >> 24) & 0x7fu;
                for (uint32_t j = 0; j < cnt; ++j) {
                    const int id = (int)its[fi + j];
                    seen[id]++;
                    for (int k = 0; k < 3; ++k)
                        if (lo[k] > b[6 * id + k] || hi[k] < b[6 * id + 3 + k]) bad_box++;
                }
            } else {
                const BvhNode& cn = nodes[l];
                for (int cc = 0; cc < OA_BVH_WIDTH; ++cc) {
                    if (cn.child[cc] == OA_BVH_EMPTY) continue;
                    if (cn.lox[cc] < lo[0] || cn.loy[cc] < lo[1] || cn.loz[cc] < lo[2] || cn.hix[cc] > hi[0] ||
                        cn.hiy[cc] > hi[1] || cn.hiz[cc] > hi[2]) bad_box++;
                }
                todo.push_back(l);
            }
        }
    }
    int coverage = 0;
    for (int i = 0; i < n; ++i) coverage += seen[i] != 1;
    std::mt19937 rng(5);
    std::normal_distribution<double> nd(0, 1);
    double ext = 0;
    for (int i = 0; i < 6 * n; ++i) ext = std::max(ext, std::fabs(b[i]));
    std::uniform_real_distribution<double> ud(-ext, ext);
    long missed = 0, visits = 0;
    size_t max_stack = 0;
    const float tmax = (float)ext;
    for (int r = 0; r < n_rays; ++r) {
        double o[3] = { ud(rng), ud(rng), ud(rng) }, d[3] = { nd(rng), nd(rng), nd(rng) };
        if (r % 7 == 0) d[r % 3] = 0.0;            /* axis-parallel planes: the clamped reciprocal */
        const double L = std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
        for (int k = 0; k < 3; ++k) d[k] /= L;
        float inv[3], oi[3];
        for (int k = 0; k < 3; ++k) {
            const float dk = (float)d[k], e = 1e-20f;
            inv[k] = 1.f / (std::fabs(dk) < e ? std::copysign(e, dk) : dk);
            oi[k] = (float)o[k] * inv[k];
        }
        auto enter = [&](const float* lo, const float* hi) {
            float tn = 0.f, tf = tmax;
            for (int k = 0; k < 3; ++k) {
                const float a = std::fma(lo[k], inv[k], -oi[k]), c = std::fma(hi[k], inv[k], -oi[k]);
                tn = std::fmax(tn, std::fmin(a, c));
                tf = std::fmin(tf, std::fmax(a, c));
            }
            return tn <= tf ? tn : INFINITY;
        };
        std::set<int> want, got;
        for (int i = 0; i < n; ++i) {
            double tn = 0, tf = tmax;
            bool ok = true;
            for (int k = 0; k < 3; ++k) {
                const double lo = b[6 * i + k] + 1e-3, hi = b[6 * i + 3 + k] - 1e-3;
                if (d[k] == 0.0) { if (o[k] < lo || o[k] > hi) ok = false; continue; }
                const double a = (lo - o[k]) / d[k], c = (hi - o[k]) / d[k];
                tn = std::max(tn, std::min(a, c));
                tf = std::min(tf, std::max(a, c));
            }
            if (ok && tn <= tf) want.insert(i);
        }
        std::vector<uint32_t> stk;
        uint32_t cur = root;
        for (;;) {
            if (cur == OA_BVH_EMPTY) {
                if (stk.empty()) break;
                cur = stk.back();
                stk.pop_back();
                continue;
            }
            if (cur & OA_BVH_LEAF_BIT) {
                const uint32_t fi = cur & 0xffffffu, cnt = (cur >> 24) & 0x7fu;
                for (uint32_t j = 0; j < cnt; ++j) got.insert((int)its[fi + j]);
                cur = OA_BVH_EMPTY;
                continue;
            }
            visits++;
            const BvhNode& nn = nodes[cur];
            float t[OA_BVH_WIDTH];
            uint32_t c[OA_BVH_WIDTH];
            for (int k = 0; k < OA_BVH_WIDTH; ++k) {
                const float lo[3] = { nn.lox[k], nn.loy[k], nn.loz[k] }, hi[3] = { nn.hix[k], nn.hiy[k], nn.hiz[k] };
                t[k] = enter(lo, hi);
                c[k] = nn.child[k];
            }
            for (int i = 1; i < OA_BVH_WIDTH; ++i)
                for (int j = i; j > 0 && t[j] < t[j - 1]; --j) { std::swap(t[j], t[j - 1]); std::swap(c[j], c[j - 1]); }
            for (int k = OA_BVH_WIDTH - 1; k >= 1; --k)
                if (t[k] < INFINITY) stk.push_back(c[k]);
            cur = t[0] < INFINITY ? c[0] : OA_BVH_EMPTY;
            max_stack = std::max(max_stack, stk.size());
        }
        for (int i : want) missed += got.count(i) == 0;
    }
    std::printf("items %d nodes %zu depth %u stack_need %u coverage %d bad_box %ld missed %ld visits %.2f max_stack %zu limit %d\n",
                n, nodes.size(), B.max_depth, B.stack_need, coverage, bad_box, missed,
                n_rays ? (double)visits / n_rays : 0.0, max_stack, OA_STACK_DEPTH);
    return 0;
}
