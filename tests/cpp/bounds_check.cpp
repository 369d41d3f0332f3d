// Host bounds of instances (opticks_amd/csrc/oa_bounds.h), the boxes the TLAS is
// built over: reads the CSGFoundry-layout arrays written by tests/test_bounds_host.py
// and writes one unpadded world box (lo xyz, hi xyz; doubles) per instance, then the
// front-instance mask (one double 0/1 per instance) and the free ball of those boxes
// (centre xyz, r^2).
#include <cstdio>
#include <cstdint>
#include <vector>
#include "oa_bounds.h"

template <typename T>
static bool read_vec(FILE* f, std::vector<T>& v)
{
    int64_t n = 0;
    if (fread(&n, sizeof(n), 1, f) != 1 || n < 0) return false;
    v.resize((size_t)n);
    return n == 0 || fread(v.data(), sizeof(T), (size_t)n, f) == (size_t)n;
}

int main(int argc, char** argv)
{
    if (argc != 3) { fprintf(stderr, "usage: bounds_check in.bin out.bin\n"); return 2; }
    FILE* f = fopen(argv[1], "rb");
    if (!f) return 2;
    std::vector<float> node, itra, prim, inst;
    std::vector<int32_t> solid;                    // (first prim, prim count) per solid
    if (!read_vec(f, node) || !read_vec(f, itra) || !read_vec(f, prim) || !read_vec(f, inst) || !read_vec(f, solid))
        return 3;
    fclose(f);
    const int64_t n_inst = (int64_t)inst.size() / 16, n_tran = (int64_t)itra.size() / 16;
    std::vector<double> out;
    std::vector<oa_bounds::Box> boxes;
    for (int64_t i = 0; i < n_inst; ++i) {
        const float* m = inst.data() + 16 * i;
        int32_t gas;
        std::memcpy(&gas, m + 7, 4);
        const oa_bounds::Box b = oa_bounds::instance_box(m, node.data(), itra.empty() ? nullptr : itra.data(), n_tran,
                                                         prim.data(), solid[2 * gas], solid[2 * gas + 1]);
        for (int k = 0; k < 3; ++k) out.push_back(b.lo[k]);
        for (int k = 0; k < 3; ++k) out.push_back(b.hi[k]);
        boxes.push_back(b);
    }
    const std::vector<char> front = oa_bounds::front_mask(boxes, 4);
    for (char f : front) out.push_back(f ? 1.0 : 0.0);
    static const float eye[16] = { 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1 };
    std::vector<oa_bounds::Box> locals;
    for (int64_t i = 0; i < n_inst; ++i) {
        int32_t gas;
        std::memcpy(&gas, inst.data() + 16 * i + 7, 4);
        locals.push_back(oa_bounds::instance_box(eye, node.data(), itra.empty() ? nullptr : itra.data(), n_tran,
                                                 prim.data(), solid[2 * gas], solid[2 * gas + 1]));
    }
    const oa_bounds::Ball ball = oa_bounds::free_ball(boxes, front, &locals, inst.data());
    for (int k = 0; k < 3; ++k) out.push_back(ball.c[k]);
    out.push_back(ball.r2);
    FILE* o = fopen(argv[2], "wb");
    if (!o) return 4;
    fwrite(out.data(), sizeof(double), out.size(), o);
    fclose(o);
    return 0;
}
