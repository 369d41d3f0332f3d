/*
 * oa_capi.cpp : the C-ABI runtime (include/opticks_amd.h).
 *
 * mi_sim_create uploads the CSGFoundry-layout geometry and the boundary
 * tables once (the QSim::UploadComponents + CSGFoundry::upload role,
 * qudarap/QSim.cc:119, CSG/CSGFoundry.cc:3392), builds one BLAS per solid and
 * a TLAS over instances (the GAS/IAS role of CSGOptiX/SBT.cc), and owns one HIP
 * stream. mi_sim_simulate* play QSim::simulate (qudarap/QSim.cc:436-620):
 * genstep prefix (seeding), genstep slicing at max_slot photons with the
 * absolute photon index carried into each slice (photon_slot_offset), the
 * propagation launch and the stable hit compaction.
 */
#include <hip/hip_runtime.h>
#include <string>
#include <vector>
#include <cstring>
#include <cstdio>
#include <cmath>
#include <algorithm>
#include <array>
#include <map>
#include <utility>

#include "../../include/opticks_amd.h"
#include "oa_layout.h"
#include "oa_portable.h"
#include "oa_bvh.h"
#include "oa_bounds.h"
#include "oa_kernels.h"
#include "oa_merge.h"

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg)
{
    g_err = msg;
    return code;
}

#define HIPCHECK(expr)                                                                          \
    do {                                                                                        \
        hipError_t e_ = (expr);                                                                 \
        if (e_ != hipSuccess) return fail(MI_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

template <typename T>
struct DBuf {
    T* p = nullptr;
    size_t n = 0;
    hipError_t reserve(size_t want)
    {
        if (want <= n && p) return hipSuccess;
        if (p) { hipError_t e = hipFree(p); if (e != hipSuccess) return e; p = nullptr; n = 0; }
        size_t bytes = std::max<size_t>(want, 1) * sizeof(T);
        hipError_t e = hipMalloc((void**)&p, bytes);
        if (e == hipSuccess) n = std::max<size_t>(want, 1);
        return e;
    }
    void release() { if (p) (void)hipFree(p); p = nullptr; n = 0; }
    void swap(DBuf& o) { std::swap(p, o.p); std::swap(n, o.n); }
};

template <typename T>
hipError_t upload(DBuf<T>& b, const T* src, size_t n, hipStream_t s)
{
    hipError_t e = b.reserve(n);
    if (e != hipSuccess) return e;
    if (n) e = hipMemcpyAsync(b.p, src, n * sizeof(T), hipMemcpyHostToDevice, s);
    if (e != hipSuccess) return e;
    return hipStreamSynchronize(s);
}

/* affine inverse in double, rounded once: same cofactor formula as the CPU checker */
void invert_affine16(const float* m, float* out16)
{
    double a[3][3], t[3], inv[3][3];
    for (int r = 0; r < 3; ++r) for (int c = 0; c < 3; ++c) a[r][c] = (double)m[4 * r + c];
    for (int c = 0; c < 3; ++c) t[c] = (double)m[12 + c];
    double c00 = a[1][1] * a[2][2] - a[1][2] * a[2][1];
    double c01 = a[1][2] * a[2][0] - a[1][0] * a[2][2];
    double c02 = a[1][0] * a[2][1] - a[1][1] * a[2][0];
    double det = a[0][0] * c00 + a[0][1] * c01 + a[0][2] * c02;
    inv[0][0] = c00 / det;
    inv[0][1] = (a[0][2] * a[2][1] - a[0][1] * a[2][2]) / det;
    inv[0][2] = (a[0][1] * a[1][2] - a[0][2] * a[1][1]) / det;
    inv[1][0] = c01 / det;
    inv[1][1] = (a[0][0] * a[2][2] - a[0][2] * a[2][0]) / det;
    inv[1][2] = (a[0][2] * a[1][0] - a[0][0] * a[1][2]) / det;
    inv[2][0] = c02 / det;
    inv[2][1] = (a[0][1] * a[2][0] - a[0][0] * a[2][1]) / det;
    inv[2][2] = (a[0][0] * a[1][1] - a[0][1] * a[1][0]) / det;
    for (int r = 0; r < 3; ++r) {
        for (int c = 0; c < 3; ++c) out16[4 * r + c] = (float)inv[r][c];
        out16[4 * r + 3] = 0.f;
    }
    for (int c = 0; c < 3; ++c) out16[12 + c] = (float)(-(t[0] * inv[0][c] + t[1] * inv[1][c] + t[2] * inv[2][c]));
    out16[15] = 1.f;
}

inline uint32_t u32_of(float f) { uint32_t u; std::memcpy(&u, &f, 4); return u; }
inline float f32_of(uint32_t u) { float f; std::memcpy(&f, &u, 4); return f; }

int tree_height(uint32_t n)
{
    int h = 0;
    while (((1u << (h + 1)) - 1u) < n) ++h;
    return ((1u << (h + 1)) - 1u) == n ? h : -1;
}

} // namespace

struct mi_sim {
    mi_config cfg;
    int device = 0;
    int n_cu = 256;
    hipStream_t stream = nullptr;
    hipEvent_t ev[4] = { nullptr, nullptr, nullptr, nullptr };
    double last_ms[4] = { 0, 0, 0, 0 };
    double adj_ms[3] = { 0, 0, 0 };

    /* geometry */
    int64_t n_inst = 0, n_prim = 0, n_tlas = 0;
    uint32_t tlas_depth = 0, blas_depth = 0, blas_stack = 0;
    uint32_t front_leaf = OA_BVH_EMPTY;             /* items of the front instances (build_tlas) */
    int32_t tlas_empty = 0;                         /* every instance is a front instance */
    std::vector<char> h_is_front;                   /* per instance: kept out of the TLAS */
    std::vector<uint32_t> solid_blas_root;
    std::vector<uint32_t> h_inst_blas_root;
    std::vector<std::array<double, 6>> solid_box;   /* padded solid-frame prim box per solid */
    std::vector<float> h_o2w;                       /* (n_inst,16) current instance transforms */
    /* host copies for the TLAS leaf boxes: CSGNode, itra and CSGPrim arrays, prim range of each solid */
    std::vector<float> h_node, h_itra, h_prim;
    std::vector<std::pair<int32_t, int32_t>> h_solid_prims;
    uint64_t generation = 0;                        /* bumped when a device buffer may move */
    int32_t ext_leaves = 0;                         /* geometry uses the extended leaf set */
    DBuf<float4> node, itra, plan;
    DBuf<InstRec> inst;
    DBuf<PrimRec> prim, blas_prims;                 /* by prim index; in BLAS leaf order */
    DBuf<BvhNode> tlas, blas;
    DBuf<uint32_t> tlas_items;
    DBuf<float> inst_box;
    DBuf<float4> free_ball;                         /* see free_ball() */

    /* physics */
    DBuf<float4> bnd;
    DBuf<int32_t> optical;
    DBuf<float> icdf;
    int32_t n_bnd = 0, n_wl = 0, n_icdf = 0, n_icdf_layer = 0, icdf_hd = 0;
    float wl_low = 0.f, wl_step = 1.f;
    DBuf<float> precooked;
    int64_t pre_ni = 0, pre_nk = 0;                 /* shape of the device copy of precooked */
    /* xorwow jump matrices: per-launch table (n_event x 800 words) and a host cache by skip */
    DBuf<uint32_t> xw_jump;
    std::map<uint64_t, std::vector<uint32_t>> xw_cache;
    /* no-allocation entry point: this GPU's scratch array (device address) and
       host workspace for an xorwow jump matrix, both set up at create */
    void* scratch = nullptr;
    std::vector<uint32_t> xw_host_J, xw_host_work;

    /* work */
    DBuf<float> gs;
    DBuf<int64_t> prefix, block_total, evbuf;
    DBuf<int32_t> word_prefix, ev_id, gs_event;
    DBuf<uint32_t> seed;                 /* genstep of each photon of the launch */
    DBuf<unsigned long long> words, next;
    /* outputs of the host-array entry point */
    DBuf<float4> photon, hit;
    DBuf<uint64_t> seq;
    std::vector<float> h_photon, h_hit, h_simtrace;
    std::vector<uint64_t> h_seq;
    /* adjoint tape of the last wavefront forward made with cfg.adjoint_tape > 0, and
       what it was made for: an adjoint of the same event, photons and gensteps uses it */
    DBuf<uint32_t> tape, tape_flag;
    DBuf<uint8_t> tape_n;
    struct TapeKey {
        bool valid = false;
        int32_t event_id = 0, k = 0;
        int64_t photon_offset = 0, n_photon = 0, n_gs = 0;
        const void* gs = nullptr;
        uint64_t generation = 0;
    } tape_key;
    /* adjoint */
    DBuf<int32_t> adj_mat_block, adj_sample_bin;
    DBuf<float> adj_rows, adj_comp;
    DBuf<uint32_t> adj_key, adj_val, adj_key2, adj_val2;
    DBuf<unsigned char> adj_tmp;
    DBuf<int64_t> adj_seg;
    /* hit merging */
    DBuf<uint8_t> m_flags;
    DBuf<uint32_t> m_sel, m_vals, m_vals2;
    DBuf<uint64_t> m_keys, m_keys2;
    DBuf<int32_t> m_len, m_off, m_counts;
    DBuf<unsigned char> m_tmp;
};

extern "C" {

void mi_config_default(mi_config* c)
{
    std::memset(c, 0, sizeof(*c));
    c->max_bounce = 31;
    c->max_time = 1e27f;
    c->tmin = 0.05f;
    c->tmin0 = 0.05f;
    c->eps0_mask = 0x4u | 0x1u | 0x2u | 0x20u | 0x10u;   /* TO,CK,SI,SC,RE */
    c->tmax = 1e6f;
    c->hit_mask = 0x40u;                                 /* SD */
    c->seed = 0;
    c->rng_offset = 0;
    c->event_skipahead = 100000;
    c->rng_kind = MI_RNG_PHILOX;
    c->output_mask = MI_OUT_PHOTON | MI_OUT_SEQ | MI_OUT_HIT;
    c->max_slot = 0;
    c->kernel = MI_KERNEL_WAVEFRONT;
    c->block_size = 64;
    c->refill_period = 1;
    c->bvh_leaf_size = 1;
    c->waves_per_cu = 0;
    c->trace_regroup_pct = 0;
    c->refill_batch = 16;
    c->adjoint_tape = 0;
}

const char* mi_last_error(void) { return g_err.c_str(); }
const char* mi_version(void) { return "opticks_amd 0.1 gfx950"; }

static int check_config(const mi_config* c)
{
    if (c->block_size != 64 && c->block_size != 128 && c->block_size != 256)
        return fail(MI_ERR_ARG, "block_size must be 64, 128 or 256");
    if (c->bvh_leaf_size < 1 || c->bvh_leaf_size > 64) return fail(MI_ERR_ARG, "bvh_leaf_size must be 1..64");
    if (c->trace_regroup_pct < 0 || c->trace_regroup_pct > 100) return fail(MI_ERR_ARG, "trace_regroup_pct must be 0..100");
    if (c->refill_batch < 1 || c->refill_batch > 64) return fail(MI_ERR_ARG, "refill_batch must be 1..64");
    if (c->adjoint_tape < 0 || c->adjoint_tape > 64) return fail(MI_ERR_ARG, "adjoint_tape must be 0..64");
    if (c->max_bounce < 0) return fail(MI_ERR_ARG, "max_bounce < 0");
    if (c->kernel != MI_KERNEL_PLAIN && c->kernel != MI_KERNEL_WAVEFRONT) return fail(MI_ERR_ARG, "unknown kernel");
    if (c->rng_kind != MI_RNG_PHILOX && c->rng_kind != MI_RNG_PRECOOKED && c->rng_kind != MI_RNG_XORWOW)
        return fail(MI_ERR_ARG, "unknown rng_kind");
    if (c->rng_kind == MI_RNG_PRECOOKED &&
        (c->precooked_ni <= 0 || c->precooked_nk <= 0 || c->precooked_nk > (int64_t)INT32_MAX))
        return fail(MI_ERR_ARG, "precooked rng needs a (ni, nk) array with ni >= 1 and 1 <= nk <= 2^31-1");
    return MI_OK;
}

static int validate_geometry(const mi_geometry* g, const mi_physics* ph)
{
    if (!g || !ph) return fail(MI_ERR_ARG, "null geometry or physics");
    if (g->n_inst <= 0 || g->n_solid <= 0 || g->n_prim <= 0 || g->n_node <= 0)
        return fail(MI_ERR_GEOMETRY, "empty geometry");
    if (g->n_inst > 0xffff) return fail(MI_ERR_GEOMETRY, "more than 65535 instances (iindex is 16 bits)");
    const unsigned char* sb = (const unsigned char*)g->solid;
    for (int64_t s = 0; s < g->n_solid; ++s) {
        int32_t np, po;
        std::memcpy(&np, sb + 48 * s + 16, 4);
        std::memcpy(&po, sb + 48 * s + 20, 4);
        if (np < 1 || po < 0 || po + np > g->n_prim) return fail(MI_ERR_GEOMETRY, "solid prim range out of bounds");
    }
    for (int64_t i = 0; i < g->n_inst; ++i) {
        int32_t gas;
        std::memcpy(&gas, g->inst + 16 * i + 7, 4);
        if (gas < 0 || gas >= g->n_solid) return fail(MI_ERR_GEOMETRY, "instance gas index out of range");
    }
    int32_t max_b = -1;
    for (int64_t p = 0; p < g->n_prim; ++p) {
        const int32_t* pr = (const int32_t*)(g->prim + 16 * p);
        int32_t num = pr[0], off = pr[1];
        if (num < 1 || off < 0 || off + num > g->n_node) return fail(MI_ERR_GEOMETRY, "prim node range out of bounds");
        const uint32_t* root = (const uint32_t*)(g->node + 16 * off);
        uint32_t tc = root[14];
        if (tc >= 101) {
            if (num != 1) return fail(MI_ERR_GEOMETRY, "leaf prim with numNode != 1");
        } else if (tc >= 1 && tc <= 3) {
            int h = tree_height(root[0]);
            if (h < 0 || (int)root[0] > num) return fail(MI_ERR_GEOMETRY, "tree subNum is not a complete tree size");
            if (h > OA_CSG_MAX_HEIGHT) return fail(MI_ERR_UNSUPPORTED, "CSG tree deeper than the device stacks");
        } else if (!(tc >= 11 && tc <= 13)) {
            return fail(MI_ERR_UNSUPPORTED, "unsupported CSG prim root typecode " + std::to_string(tc));
        }
        const int tree_nodes = (tc >= 1 && tc <= 3) ? (int)root[0] : 1;
        for (int k = 0; k < num; ++k) {
            const uint32_t* nd = (const uint32_t*)(g->node + 16 * (off + k));
            uint32_t t = nd[14];
            /* sphere, zsphere, cylinder, cone, box3, convexpolyhedron, hyperboloid, phicut, halfspace */
            const bool leaf = t == 101 || t == 103 || t == 105 || t == 108 || t == 110 || t == 112 || t == 117 ||
                              t == 121 || t == 125;
            const bool list = t >= 11 && t <= 13;
            bool ok = k < tree_nodes ? (t == 0 || (t >= 1 && t <= 3) || leaf || list) : leaf;
            if (!ok) return fail(MI_ERR_UNSUPPORTED, "unsupported CSG node typecode " + std::to_string(t));
            if (list) {
                /* sub leaves at root + subOffset (CSG/CSGNode.h:187-211), after the tree nodes */
                const uint64_t sn = nd[0], so = nd[1];
                if (sn < 1 || so < (uint64_t)tree_nodes || so + sn > (uint64_t)num)
                    return fail(MI_ERR_GEOMETRY, "list node sub range outside its prim");
                if (t == 11 && sn > 8)
                    return fail(MI_ERR_UNSUPPORTED, "CSG_CONTIGUOUS with more than 8 leaves (the reference's enter arrays)");
            }
            if (t == 112) {
                const uint64_t pidx = nd[0], pnum = nd[1];
                if (!g->plan || pidx + pnum > (uint64_t)g->n_plan)
                    return fail(MI_ERR_GEOMETRY, "convexpolyhedron planes beyond the plan array");
            }
            uint32_t gt = nd[15] & 0x7fffffffu;
            if (gt > (uint32_t)g->n_tran) return fail(MI_ERR_GEOMETRY, "gtransformIdx beyond itra");
            if (gt && !g->itra) return fail(MI_ERR_GEOMETRY, "node transform without itra array");
        }
        max_b = std::max<int32_t>(max_b, (int32_t)root[6]);
    }
    if (max_b >= ph->n_bnd) return fail(MI_ERR_GEOMETRY, "prim boundary beyond the bnd table");
    if (!ph->bnd || !ph->optical || ph->n_wl < 2) return fail(MI_ERR_ARG, "physics tables missing");
    return MI_OK;
}

/* padded world boxes of the instances from the current transforms: per prim of
   the instance's solid, the exact world bound of its CSG leaves (oa_bounds.h)
   intersected with the transformed corners of its CSGPrim AABB, joined over the
   solid's prims */
static void instance_items(const mi_sim* sim, std::vector<oa_bvh::Item>& titems)
{
    const int64_t ni = sim->n_inst;
    const int64_t n_tran = (int64_t)sim->h_itra.size() / 16;
    titems.resize(ni);
    for (int64_t i = 0; i < ni; ++i) {
        const float* m = sim->h_o2w.data() + 16 * i;
        int32_t gas;
        std::memcpy(&gas, m + 7, 4);
        const std::pair<int32_t, int32_t> pr = sim->h_solid_prims[gas];
        const oa_bounds::Box all = oa_bounds::instance_box(m, sim->h_node.data(),
                                                           sim->h_itra.empty() ? nullptr : sim->h_itra.data(), n_tran,
                                                           sim->h_prim.data(), pr.first, pr.second);
        oa_bvh::Item& it = titems[i];
        for (int k = 0; k < 3; ++k) {
            it.lo[k] = all.lo[k]; it.hi[k] = all.hi[k];
            if (!(it.hi[k] >= it.lo[k])) { it.lo[k] = it.hi[k] = 0.5 * (all.lo[k] + all.hi[k]); }
            const double pad = 1e-3 * (it.hi[k] - it.lo[k]) + 1e-2;
            it.lo[k] -= pad; it.hi[k] += pad;
        }
        it.id = (uint32_t)i;
    }
}

static oa_bounds::Box item_box(const oa_bvh::Item& it)
{
    oa_bounds::Box b;
    for (int k = 0; k < 3; ++k) { b.lo[k] = it.lo[k]; b.hi[k] = it.hi[k]; }
    return b;
}

/*
 * Front instances (oa_bounds::front_mask: world-sized boxes, at most OA_MAX_FRONT)
 * stay out of the TLAS; every trace tests them first (their items follow the TLAS
 * items, reached through SimParams.front_leaf), so the whole wave runs them
 * together and the TLAS walk starts with the nearest of their hits as its tmax.
 * The nearest hit is chosen by (t, instance, prim), not by test order, so which
 * instances are front changes no result.
 */
static void split_front(const std::vector<oa_bvh::Item>& all, std::vector<oa_bvh::Item>& rest,
                        std::vector<uint32_t>& front, std::vector<char>& is_front)
{
    std::vector<oa_bounds::Box> boxes(all.size());
    for (size_t i = 0; i < all.size(); ++i) boxes[i] = item_box(all[i]);
    is_front = oa_bounds::front_mask(boxes, OA_MAX_FRONT);
    front.clear();
    rest.clear();
    for (size_t i = 0; i < all.size(); ++i) {
        if (is_front[i]) front.push_back(all[i].id);
        else rest.push_back(all[i]);
    }
}

/*
 * The free ball (oa_bounds::free_ball over the padded boxes the TLAS is built
 * over): a trace whose segment from its origin to the nearest front-instance hit
 * has both ends inside it cannot hit a TLAS instance (the ball is convex), so that
 * lane skips the TLAS walk -- on a detector shell around a scintillator volume,
 * every trace that starts inside the shell and ends on an inner surface; a lane
 * whose origin is inside it starts its box tests where the ray leaves it. It culls
 * only, so no result depends on it.
 */
static float4 free_ball(const mi_sim* sim, const std::vector<oa_bvh::Item>& all, const std::vector<char>& is_front)
{
    std::vector<oa_bounds::Box> boxes(all.size()), locals(all.size());
    for (size_t i = 0; i < all.size(); ++i) boxes[i] = item_box(all[i]);
    /* each solid's box in its own frame (the exact leaf bounds under the identity) */
    static const float eye[16] = { 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1 };
    const int64_t n_tran = (int64_t)sim->h_itra.size() / 16;
    std::vector<oa_bounds::Box> solid_local(sim->h_solid_prims.size());
    for (size_t sidx = 0; sidx < solid_local.size(); ++sidx)
        solid_local[sidx] = oa_bounds::instance_box(eye, sim->h_node.data(),
                                                    sim->h_itra.empty() ? nullptr : sim->h_itra.data(), n_tran,
                                                    sim->h_prim.data(), sim->h_solid_prims[sidx].first,
                                                    sim->h_solid_prims[sidx].second);
    for (size_t i = 0; i < all.size(); ++i) {
        int32_t gas;
        std::memcpy(&gas, sim->h_o2w.data() + 16 * i + 7, 4);
        locals[i] = solid_local[gas];
    }
    const oa_bounds::Ball b = oa_bounds::free_ball(boxes, is_front, &locals, sim->h_o2w.data());
    return make_float4((float)b.c[0], (float)b.c[1], (float)b.c[2], (float)b.r2);
}

/* host SAH build of the TLAS over instance boxes with the given leaf size, uploaded
   into new buffers that replace the handle's only when every step succeeded */
static int build_tlas(mi_sim* sim, int leaf)
{
    std::vector<oa_bvh::Item> all, titems;
    std::vector<uint32_t> front;
    instance_items(sim, all);
    std::vector<char> is_front;
    split_front(all, titems, front, is_front);
    const float4 ball = free_ball(sim, all, is_front);
    std::vector<BvhNode> tlas;
    std::vector<uint32_t> tlas_items;
    oa_bvh::Builder tb(tlas, tlas_items, leaf);
    uint32_t troot = tb.build(titems);
    if (troot != 0) return fail(MI_ERR_GEOMETRY, "TLAS root is not node 0");
    uint32_t front_leaf = OA_BVH_EMPTY;
    if (!front.empty()) {
        front_leaf = OA_BVH_LEAF_BIT | ((uint32_t)front.size() << 24) | (uint32_t)tlas_items.size();
        tlas_items.insert(tlas_items.end(), front.begin(), front.end());
    }
    /* deepest stack a walk can reach: TLAS pushes, then the BLAS walk on top */
    if (tb.stack_need + sim->blas_stack + 1 > OA_STACK_DEPTH)
        return fail(MI_ERR_UNSUPPORTED, "BVH deeper than the LDS traversal stack");
    DBuf<BvhNode> nt;
    DBuf<uint32_t> ni;
    DBuf<float4> nb;
    hipError_t e = hipStreamSynchronize(sim->stream);
    if (e == hipSuccess) e = upload(nt, tlas.data(), tlas.size(), sim->stream);
    if (e == hipSuccess) e = upload(ni, tlas_items.data(), tlas_items.size(), sim->stream);
    if (e == hipSuccess) e = upload(nb, &ball, 1, sim->stream);
    if (e != hipSuccess) {
        nt.release(); ni.release(); nb.release();
        return fail(MI_ERR_HIP, std::string("TLAS upload: ") + hipGetErrorString(e));
    }
    sim->tlas.swap(nt);
    sim->tlas_items.swap(ni);
    sim->free_ball.swap(nb);
    nt.release(); ni.release(); nb.release();
    sim->h_is_front = is_front;
    sim->tlas_depth = tb.max_depth;
    sim->n_tlas = (int64_t)tlas.size();
    sim->front_leaf = front_leaf;
    sim->tlas_empty = titems.empty() ? 1 : 0;
    sim->generation++;
    return MI_OK;
}

static int build_and_upload_geometry(mi_sim* sim, const mi_geometry* g)
{
    hipStream_t s = sim->stream;
    const int64_t ni = g->n_inst;
    /* BLAS per solid over its prims (solid frame boxes from CSGPrim q2/q3) */
    std::vector<BvhNode> blas;
    std::vector<uint32_t> blas_items;
    const unsigned char* sb = (const unsigned char*)g->solid;
    sim->solid_blas_root.resize(g->n_solid);
    uint32_t bdepth = 0, bneed = 0;
    std::vector<std::array<double, 6>> solid_box(g->n_solid);
    std::vector<uint32_t> solid_flat(g->n_solid);      /* one leaf link over all the solid's prims */
    for (int64_t so = 0; so < g->n_solid; ++so) {
        int32_t np, po;
        std::memcpy(&np, sb + 48 * so + 16, 4);
        std::memcpy(&po, sb + 48 * so + 20, 4);
        std::vector<oa_bvh::Item> items(np);
        std::array<double, 6> bx = { 1e300, 1e300, 1e300, -1e300, -1e300, -1e300 };
        for (int p = 0; p < np; ++p) {
            const float* pr = g->prim + 16 * (po + p);
            oa_bvh::Item& it = items[p];
            double lo[3] = { pr[8], pr[9], pr[10] }, hi[3] = { pr[11], pr[12], pr[13] };
            for (int k = 0; k < 3; ++k) {
                double pad = 1e-3 * (hi[k] - lo[k]) + 1e-2;
                it.lo[k] = lo[k] - pad; it.hi[k] = hi[k] + pad;
                bx[k] = std::min(bx[k], it.lo[k]); bx[3 + k] = std::max(bx[3 + k], it.hi[k]);
            }
            it.id = (uint32_t)(po + p);
        }
        oa_bvh::Builder b(blas, blas_items, 1);
        const uint32_t start = (uint32_t)blas_items.size();
        sim->solid_blas_root[so] = b.build(items);
        /* the leaves of a solid's BLAS hold each of its prims once, appended together */
        solid_flat[so] = (np > 0 && np <= 0x7f && (size_t)start + np <= 0x00ffffffu)
                             ? OA_BVH_LEAF_BIT | ((uint32_t)np << 24) | start : sim->solid_blas_root[so];
        bdepth = std::max(bdepth, b.max_depth);
        bneed = std::max(bneed, b.stack_need);
        solid_box[so] = bx;
    }
    /* instances: world->object transforms */
    std::vector<InstRec> inst(ni);
    sim->h_inst_blas_root.resize(ni);
    sim->solid_box = solid_box;
    sim->h_o2w.assign(g->inst, g->inst + 16 * ni);
    sim->h_node.assign(g->node, g->node + 16 * g->n_node);
    if (g->itra && g->n_tran > 0) sim->h_itra.assign(g->itra, g->itra + 16 * g->n_tran);
    sim->h_prim.assign(g->prim, g->prim + 16 * g->n_prim);
    sim->h_solid_prims.resize(g->n_solid);
    for (int64_t so = 0; so < g->n_solid; ++so) {
        int32_t np, po;
        std::memcpy(&np, sb + 48 * so + 16, 4);
        std::memcpy(&po, sb + 48 * so + 20, 4);
        sim->h_solid_prims[so] = { po, np };
    }
    for (int64_t i = 0; i < ni; ++i) {
        const float* m = g->inst + 16 * i;
        int32_t gas, ident;
        std::memcpy(&gas, m + 7, 4);
        std::memcpy(&ident, m + 11, 4);
        InstRec& r = inst[i];
        invert_affine16(m, r.m);
        r.m[3] = f32_of((uint32_t)gas);
        r.m[7] = f32_of((uint32_t)ident);
        r.m[11] = f32_of(sim->solid_blas_root[gas]);
        sim->h_inst_blas_root[i] = sim->solid_blas_root[gas];
    }
    sim->blas_depth = bdepth;
    sim->blas_stack = bneed;
    sim->n_inst = ni;
    sim->n_prim = g->n_prim;
    int trc = build_tlas(sim, sim->cfg.bvh_leaf_size);
    if (trc) return trc;
    /* Front instances (tested by every trace, on all lanes together) walk no BLAS:
       their solid's prims are one leaf, tested in turn without box tests. Nearly
       every trace meets all of a world-sized solid's prim boxes anyway, and the
       nearest hit does not depend on the order or on box pruning (ties go to the
       smaller key). The link is valid for any instance, so a later TLAS rebuild
       that changes which instances are front changes no result. */
    for (int64_t i = 0; i < ni; ++i) {
        if (!sim->h_is_front[i]) continue;
        int32_t gas;
        std::memcpy(&gas, g->inst + 16 * i + 7, 4);
        inst[i].m[11] = f32_of(solid_flat[gas]);
    }

    std::vector<PrimRec> prim(g->n_prim);
    for (int64_t p = 0; p < g->n_prim; ++p) {
        const int32_t* pr = (const int32_t*)(g->prim + 16 * p);
        const uint32_t* root = (const uint32_t*)(g->node + 16 * pr[1]);
        prim[p].node_offset = (uint32_t)pr[1];
        prim[p].boundary = root[6];
        prim[p].global_prim = (uint32_t)pr[15];
        prim[p].index = (uint32_t)p;
    }
    /* the leaves' items replaced by the records they name, so the BLAS leaf links index them */
    std::vector<PrimRec> blas_prims(blas_items.size());
    for (size_t k = 0; k < blas_items.size(); ++k) blas_prims[k] = prim[blas_items[k]];
    sim->ext_leaves = 0;
    for (int64_t k = 0; k < g->n_node; ++k) {
        const uint32_t t = ((const uint32_t*)(g->node + 16 * k))[14];
        if (t == 108 || t == 112 || t == 117 || t == 121 || t == 125 || (t >= 11 && t <= 13)) sim->ext_leaves = 1;
    }
    HIPCHECK(upload(sim->node, (const float4*)g->node, (size_t)g->n_node * 4, s));
    if (g->n_tran > 0) HIPCHECK(upload(sim->itra, (const float4*)g->itra, (size_t)g->n_tran * 4, s));
    else HIPCHECK(sim->itra.reserve(4));
    if (g->n_plan > 0 && g->plan) HIPCHECK(upload(sim->plan, (const float4*)g->plan, (size_t)g->n_plan, s));
    else HIPCHECK(sim->plan.reserve(1));
    HIPCHECK(upload(sim->inst, inst.data(), inst.size(), s));
    HIPCHECK(upload(sim->prim, prim.data(), prim.size(), s));
    HIPCHECK(upload(sim->blas, blas.data(), blas.size(), s));
    HIPCHECK(upload(sim->blas_prims, blas_prims.data(), blas_prims.size(), s));
    HIPCHECK(sim->inst_box.reserve((size_t)ni * 6));
    return MI_OK;
}

static int upload_physics(mi_sim* sim, const mi_physics* ph)
{
    hipStream_t s = sim->stream;
    const size_t nb = (size_t)ph->n_bnd * 4 * 2 * ph->n_wl;
    HIPCHECK(upload(sim->bnd, (const float4*)ph->bnd, nb, s));
    HIPCHECK(upload(sim->optical, ph->optical, (size_t)ph->n_bnd * 16, s));
    sim->n_bnd = ph->n_bnd;
    sim->n_wl = ph->n_wl;
    sim->wl_low = ph->wl_low;
    sim->wl_step = ph->wl_step;
    if (ph->icdf && ph->n_icdf > 0 && ph->n_icdf_layer > 0) {
        HIPCHECK(upload(sim->icdf, ph->icdf, (size_t)ph->n_icdf * ph->n_icdf_layer, s));
        sim->n_icdf = ph->n_icdf;
        sim->n_icdf_layer = ph->n_icdf_layer;
        sim->icdf_hd = ph->icdf_hd_factor;
    } else {
        float one = 0.f;
        HIPCHECK(upload(sim->icdf, &one, 1, s));
        sim->n_icdf = 1;
        sim->n_icdf_layer = 1;
        sim->icdf_hd = 0;
    }
    return MI_OK;
}

/* settings frozen into a captured batch's launch parameters */
static bool same_physics(const mi_config& a, const mi_config& b)
{
    return a.max_bounce == b.max_bounce && a.max_time == b.max_time && a.tmin == b.tmin && a.tmin0 == b.tmin0 &&
           a.eps0_mask == b.eps0_mask && a.tmax == b.tmax && a.hit_mask == b.hit_mask && a.seed == b.seed &&
           a.rng_offset == b.rng_offset && a.event_skipahead == b.event_skipahead && a.rng_kind == b.rng_kind;
}

/*
 * Validate, build (TLAS for a changed leaf size, precooked upload) and only
 * then commit. The handle keeps its own device copy of the precooked array;
 * its config holds a null pointer with that copy's shape, and a config with a
 * null pointer and the same shape keeps the copy. tlas_built: the TLAS of
 * cfg's leaf size is already in place (create).
 */
static int apply_config(mi_sim* sim, const mi_config* cfg, bool tlas_built)
{
    int rc = check_config(cfg);
    if (rc) return rc;
    const bool pre = cfg->rng_kind == MI_RNG_PRECOOKED;
    if (pre && !cfg->precooked &&
        !(sim->precooked.p && sim->pre_ni == cfg->precooked_ni && sim->pre_nk == cfg->precooked_nk))
        return fail(MI_ERR_ARG, "precooked rng needs the (ni, nk) array (or null with the shape the handle holds)");
    DBuf<float> npre;
    if (pre && cfg->precooked) {
        hipError_t e = upload(npre, cfg->precooked, (size_t)(cfg->precooked_ni * cfg->precooked_nk), sim->stream);
        if (e != hipSuccess) {
            npre.release();
            return fail(MI_ERR_HIP, std::string("precooked upload: ") + hipGetErrorString(e));
        }
    }
    if (!tlas_built && cfg->bvh_leaf_size != sim->cfg.bvh_leaf_size) {
        rc = build_tlas(sim, cfg->bvh_leaf_size);   /* bumps generation */
        if (rc) { npre.release(); return rc; }
    }
    if (npre.p) {
        sim->precooked.swap(npre);
        npre.release();
        sim->pre_ni = cfg->precooked_ni;
        sim->pre_nk = cfg->precooked_nk;
        sim->generation++;
    }
    if (!same_physics(sim->cfg, *cfg)) sim->generation++;
    sim->cfg = *cfg;
    sim->cfg.precooked = nullptr;
    if (!pre) { sim->cfg.precooked_ni = sim->pre_ni; sim->cfg.precooked_nk = sim->pre_nk; }
    return MI_OK;
}

int mi_sim_create(const mi_geometry* geo, const mi_physics* phys, const mi_config* cfg, mi_sim** out)
{
    if (!out) return fail(MI_ERR_ARG, "null out");
    *out = nullptr;
    mi_config c;
    if (cfg) c = *cfg; else mi_config_default(&c);
    int rc = check_config(&c);
    if (rc) return rc;
    rc = validate_geometry(geo, phys);
    if (rc) return rc;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return fail(MI_ERR_HIP, "no HIP device");
    mi_sim* sim = new mi_sim();
    sim->cfg = c;
    HIPCHECK(hipGetDevice(&sim->device));
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, sim->device) == hipSuccess) sim->n_cu = prop.multiProcessorCount;
    if (hipStreamCreate(&sim->stream) != hipSuccess) { delete sim; return fail(MI_ERR_HIP, "hipStreamCreate failed"); }
    for (auto& e : sim->ev) HIPCHECK(hipEventCreate(&e));
    rc = build_and_upload_geometry(sim, geo);
    if (!rc) rc = upload_physics(sim, phys);
    if (!rc) rc = apply_config(sim, &c, true);
    if (!rc) {
        rc = sim->next.reserve(1) == hipSuccess ? MI_OK : fail(MI_ERR_HIP, "alloc counter");
    }
    if (!rc) {
        hipError_t e = oa_launch::noalloc_scratch(sim->device, &sim->scratch);
        if (e != hipSuccess) rc = fail(MI_ERR_HIP, std::string("scratch address: ") + hipGetErrorString(e));
        sim->xw_host_J.resize(OA_XORWOW_JUMP_WORDS);
        sim->xw_host_work.resize(OA_XORWOW_WORK_WORDS);
    }
    if (rc) { mi_sim_destroy(sim); return rc; }
    *out = sim;
    return MI_OK;
}

void mi_sim_destroy(mi_sim* sim)
{
    if (!sim) return;
    if (sim->stream) (void)hipStreamSynchronize(sim->stream);
    sim->node.release(); sim->itra.release(); sim->plan.release(); sim->inst.release(); sim->prim.release();
    sim->tlas.release(); sim->blas.release(); sim->tlas_items.release(); sim->blas_prims.release();
    sim->inst_box.release(); sim->free_ball.release();
    sim->bnd.release(); sim->optical.release(); sim->icdf.release(); sim->precooked.release();
    sim->gs.release(); sim->prefix.release(); sim->block_total.release(); sim->evbuf.release();
    sim->word_prefix.release(); sim->ev_id.release(); sim->words.release(); sim->next.release();
    sim->seed.release(); sim->gs_event.release();
    sim->photon.release(); sim->hit.release(); sim->seq.release();
    sim->adj_mat_block.release(); sim->adj_sample_bin.release(); sim->adj_rows.release(); sim->adj_comp.release();
    sim->tape.release(); sim->tape_n.release(); sim->tape_flag.release();
    sim->adj_key.release(); sim->adj_val.release(); sim->adj_key2.release(); sim->adj_val2.release();
    sim->adj_tmp.release(); sim->adj_seg.release();
    sim->m_flags.release(); sim->m_sel.release(); sim->m_vals.release(); sim->m_vals2.release();
    sim->m_keys.release(); sim->m_keys2.release(); sim->m_len.release(); sim->m_off.release();
    sim->m_counts.release(); sim->m_tmp.release();
    for (auto& e : sim->ev) if (e) (void)hipEventDestroy(e);
    if (sim->stream) (void)hipStreamDestroy(sim->stream);
    delete sim;
}

int mi_sim_set_config(mi_sim* sim, const mi_config* cfg)
{
    if (!sim || !cfg) return fail(MI_ERR_ARG, "null");
    return apply_config(sim, cfg, false);
}

int mi_sim_get_config(const mi_sim* sim, mi_config* cfg)
{
    if (!sim || !cfg) return fail(MI_ERR_ARG, "null");
    *cfg = sim->cfg;
    return MI_OK;
}

static SimParams base_params(const mi_sim* sim)
{
    SimParams P;
    std::memset(&P, 0, sizeof(P));
    P.node = sim->node.p; P.itra = sim->itra.p; P.plan = sim->plan.p; P.ext_leaves = sim->ext_leaves; P.inst = sim->inst.p; P.prim = sim->prim.p;
    P.tlas = sim->tlas.p; P.tlas_items = sim->tlas_items.p; P.blas = sim->blas.p; P.blas_prims = sim->blas_prims.p;
    P.front_leaf = sim->front_leaf; P.tlas_empty = sim->tlas_empty; P.free_ball = sim->free_ball.p;
    P.bnd = sim->bnd.p; P.optical = sim->optical.p; P.icdf = sim->icdf.p;
    P.n_wl = sim->n_wl; P.wl_low = sim->wl_low; P.wl_step = sim->wl_step;
    P.n_icdf = sim->n_icdf; P.n_icdf_layer = sim->n_icdf_layer; P.icdf_hd = sim->icdf_hd;
    const mi_config& c = sim->cfg;
    P.max_bounce = c.max_bounce; P.max_time = c.max_time; P.tmin = c.tmin; P.tmin0 = c.tmin0; P.tmax = c.tmax;
    P.eps0_mask = c.eps0_mask; P.hit_mask = c.hit_mask;
    P.seed = c.seed; P.rng_offset = c.rng_offset; P.skipahead = c.event_skipahead;
    P.rng_kind = c.rng_kind;
    P.precooked = c.rng_kind == MI_RNG_PRECOOKED ? sim->precooked.p : nullptr;
    P.pre_ni = sim->pre_ni; P.pre_nk = sim->pre_nk;
    P.xw_jump = nullptr;
    P.next = sim->next.p;
    P.refill_period = c.refill_period;
    P.regroup_pct = c.trace_regroup_pct;
    P.refill_batch = c.refill_batch;
    return P;
}

/* every photon index of a launch, [0, end_index), needs a precooked row */
static int check_precooked_rows(const mi_sim* sim, int64_t end_index)
{
    if (sim->cfg.rng_kind == MI_RNG_PRECOOKED && end_index > sim->pre_ni)
        return fail(MI_ERR_ARG, "precooked uniforms hold " + std::to_string(sim->pre_ni) +
                                    " rows; the launch needs photon indices up to " + std::to_string(end_index - 1));
    return MI_OK;
}

/* xorwow: the jump matrix of each event's skip (offset + event_id * skipahead)
   uploaded into dst (n_event x 800 words); *table stays null for other generators */
static int xorwow_table(mi_sim* sim, const int32_t* event_ids, int32_t n_event, DBuf<uint32_t>& dst,
                        const uint32_t** table)
{
    *table = nullptr;
    if (sim->cfg.rng_kind != MI_RNG_XORWOW) return MI_OK;
    std::vector<uint32_t> h((size_t)n_event * OA_XORWOW_JUMP_WORDS);
    std::vector<uint32_t> work;
    for (int32_t e = 0; e < n_event; ++e) {
        const uint64_t n = sim->cfg.rng_offset + sim->cfg.event_skipahead * (uint64_t)(uint32_t)event_ids[e];
        auto it = sim->xw_cache.find(n);
        if (it == sim->xw_cache.end()) {
            std::vector<uint32_t> J(OA_XORWOW_JUMP_WORDS);
            work.resize(OA_XORWOW_WORK_WORDS);
            oa_xorwow_jump_matrix(n, J.data(), work.data());
            if (sim->xw_cache.size() >= 4096) sim->xw_cache.clear();
            it = sim->xw_cache.emplace(n, std::move(J)).first;
        }
        std::copy(it->second.begin(), it->second.end(), h.begin() + (size_t)e * OA_XORWOW_JUMP_WORDS);
    }
    HIPCHECK(upload(dst, h.data(), h.size(), sim->stream));
    *table = dst.p;
    return MI_OK;
}

/* persistent grid: every resident slot of every CU (occupancy API), or the
   configured waves per CU */
static int wave_grid(const mi_sim* sim)
{
    const int waves_per_block = sim->cfg.block_size / 64;
    if (sim->cfg.waves_per_cu > 0) return std::max(1, sim->n_cu * sim->cfg.waves_per_cu / waves_per_block);
    int nb = oa_launch::wave_blocks_per_cu(sim->cfg.block_size, sim->cfg.rng_kind);
    if (nb <= 0) nb = std::max(1, 8 / waves_per_block);
    return std::max(1, sim->n_cu * nb);
}

/*
 * One propagation launch over gensteps already on the device with a
 * launch-local prefix; events described by host arrays (copied to evbuf).
 * Writes photons [0,n_photon) of the launch and the hit bitmask words.
 */
static int launch(mi_sim* sim, SimParams& P, hipStream_t s)
{
    const int64_t n_words = (P.n_photon + 63) / 64;
    HIPCHECK(sim->words.reserve((size_t)n_words + 1));
    P.hitwords = sim->words.p;
    HIPCHECK(sim->seed.reserve((size_t)P.n_photon));
    HIPCHECK(sim->gs_event.reserve((size_t)std::max<int64_t>(P.n_gs, 1)));
    HIPCHECK(oa_launch::seed_map(P.gs_prefix, P.n_gs, P.n_photon, sim->seed.p, P.ev_gs_begin, P.n_event,
                                 sim->gs_event.p, s));
    P.seed_map = sim->seed.p;
    P.gs_event = sim->gs_event.p;
    if (sim->cfg.kernel == MI_KERNEL_WAVEFRONT) {
        HIPCHECK(oa_launch::zero_u64(sim->words.p, n_words, s));
        HIPCHECK(oa_launch::zero_u64(sim->next.p, 1, s));
    }
    HIPCHECK(oa_launch::simulate(P, sim->cfg.kernel, sim->cfg.block_size, wave_grid(sim), s));
    return MI_OK;
}

/* stable hit compaction; the hit count lands in block_total[nb], read after the end event */
static int compact(mi_sim* sim, int64_t n_photon, const float4* d_photon, float4* d_hit, int64_t hit_cap,
                   int64_t** d_count, hipStream_t s)
{
    const int64_t n_words = (n_photon + 63) / 64;
    const int64_t nb = (n_words + 255) / 256;
    HIPCHECK(sim->word_prefix.reserve((size_t)n_words + 1));
    HIPCHECK(sim->block_total.reserve((size_t)nb + 1));
    HIPCHECK(oa_launch::compact_hits(sim->words.p, n_words, sim->word_prefix.p, sim->block_total.p,
                                     d_photon, d_hit, hit_cap, s));
    *d_count = sim->block_total.p + nb;
    return MI_OK;
}

static int simulate_device_impl(mi_sim* sim, int32_t event_id, const void* d_gensteps, int64_t n_gs, int64_t n_photon,
                                int64_t photon_offset, void* d_photon, void* d_seq, void* d_hit, int64_t hit_capacity,
                                int64_t* n_hit_out, void* d_lite, void* stream)
{
    if (!sim) return fail(MI_ERR_ARG, "null sim");
    if (n_hit_out) *n_hit_out = 0;
    if (n_gs <= 0) return fail(MI_NOTHING_TO_DO, "zero gensteps");
    if (!d_gensteps) return fail(MI_ERR_ARG, "null gensteps");
    if (d_hit && !d_photon) return fail(MI_ERR_ARG, "hits are gathered from the photon array: pass d_photon");
    hipStream_t s = stream ? (hipStream_t)stream : sim->stream;
    HIPCHECK(hipEventRecord(sim->ev[0], s));
    HIPCHECK(sim->prefix.reserve((size_t)n_gs + 1));
    HIPCHECK(oa_launch::gs_prefix((const float*)d_gensteps, n_gs, sim->prefix.p, s));
    std::vector<int64_t> hprefix;
    const int64_t max_slot = sim->cfg.max_slot;
    int64_t total = 0;
    if (max_slot > 0) {
        hprefix.resize((size_t)n_gs + 1);
        HIPCHECK(hipMemcpyAsync(hprefix.data(), sim->prefix.p, sizeof(int64_t) * (n_gs + 1), hipMemcpyDeviceToHost, s));
        HIPCHECK(hipStreamSynchronize(s));
        total = hprefix[n_gs];
        if (n_photon >= 0 && n_photon != total)
            return fail(MI_ERR_ARG, "with max_slot slicing n_photon must equal the gensteps' photons");
    } else {
        HIPCHECK(hipMemcpyAsync(&total, sim->prefix.p + n_gs, sizeof(int64_t), hipMemcpyDeviceToHost, s));
        HIPCHECK(hipStreamSynchronize(s));
        if (n_photon > total)
            return fail(MI_ERR_ARG, "n_photon " + std::to_string(n_photon) + " exceeds the gensteps' " +
                                        std::to_string(total) + " photons");
    }
    if (n_photon < 0) n_photon = total;
    HIPCHECK(hipEventRecord(sim->ev[1], s));
    if (n_photon == 0) return fail(MI_NOTHING_TO_DO, "gensteps carry no photons");
    int prc = check_precooked_rows(sim, photon_offset + n_photon);
    if (prc) return prc;
    const uint32_t* xw_table = nullptr;
    prc = xorwow_table(sim, &event_id, 1, sim->xw_jump, &xw_table);
    if (prc) return prc;
    /* slices of whole gensteps with at most max_slot photons (sysrap/SGenstep.h:248) */
    std::vector<std::pair<int64_t, int64_t>> slices;
    if (max_slot > 0 && n_photon > max_slot) {
        int64_t g0 = 0;
        while (g0 < n_gs) {
            int64_t g1 = g0 + 1;
            while (g1 < n_gs && hprefix[g1 + 1] - hprefix[g0] <= max_slot) ++g1;
            if (hprefix[g1] - hprefix[g0] > max_slot)
                return fail(MI_ERR_CAPACITY, "one genstep has more photons than max_slot");
            slices.emplace_back(g0, g1);
            g0 = g1;
        }
    } else {
        slices.emplace_back(0, n_gs);
    }
    /* adjoint tape: wavefront kernel, full photon records (not the lite form) */
    const int32_t K = sim->cfg.adjoint_tape;
    const bool taping = K > 0 && sim->cfg.kernel == MI_KERNEL_WAVEFRONT && !d_lite;
    sim->tape_key.valid = false;
    if (taping) {
        HIPCHECK(sim->tape.reserve((size_t)n_photon * 3 * (size_t)K));
        HIPCHECK(sim->tape_n.reserve((size_t)n_photon));
    }
    int64_t hits = 0;
    double prop_ms = 0.0, comp_ms = 0.0;
    for (auto& sl : slices) {
        const int64_t g0 = sl.first, g1 = sl.second;
        const int64_t p0 = hprefix.empty() ? 0 : hprefix[g0];
        const int64_t np = hprefix.empty() ? n_photon : hprefix[g1] - hprefix[g0];
        if (slices.size() > 1) {
            HIPCHECK(oa_launch::gs_prefix((const float*)d_gensteps + 24 * g0, g1 - g0, sim->prefix.p, s));
        }
        int64_t evh[4] = { 0, g1 - g0, 0, np };
        int32_t evid = event_id;
        HIPCHECK(sim->evbuf.reserve(4));
        HIPCHECK(sim->ev_id.reserve(1));
        HIPCHECK(hipMemcpyAsync(sim->evbuf.p, evh, sizeof(evh), hipMemcpyHostToDevice, s));
        HIPCHECK(hipMemcpyAsync(sim->ev_id.p, &evid, sizeof(evid), hipMemcpyHostToDevice, s));
        SimParams P = base_params(sim);
        P.xw_jump = xw_table;
        P.gensteps = (const float*)d_gensteps + 24 * g0;
        P.gs_prefix = sim->prefix.p;
        P.n_gs = g1 - g0;
        P.n_event = 1;
        P.ev_gs_begin = sim->evbuf.p;
        P.ev_ph_begin = sim->evbuf.p + 2;
        P.ev_id = sim->ev_id.p;
        P.n_photon = np;
        P.photon_offset = photon_offset + p0;
        P.photon = d_photon ? (float4*)d_photon + 4 * p0 : nullptr;
        P.seq = d_seq ? (uint64_t*)d_seq + 4 * p0 : nullptr;
        P.photonlite = d_lite ? (uint4*)d_lite + p0 : nullptr;
        if (taping) {
            /* segment-major records with the whole launch's photon count as the stride */
            P.tape = sim->tape.p + (size_t)p0 * 3;
            P.tape_n = sim->tape_n.p + p0;
            P.tape_stride = n_photon;
            P.tape_k = K;
        }
        int rc = launch(sim, P, s);
        if (rc) return rc;
        HIPCHECK(hipEventRecord(sim->ev[2], s));
        int64_t nh = 0;
        int64_t* d_count = nullptr;
        rc = compact(sim, np, P.photon, d_hit ? (float4*)d_hit + 4 * hits : nullptr,
                     d_hit ? std::max<int64_t>(0, hit_capacity - hits) : 0, &d_count, s);
        if (rc) return rc;
        HIPCHECK(hipEventRecord(sim->ev[3], s));
        HIPCHECK(hipMemcpyAsync(&nh, d_count, sizeof(int64_t), hipMemcpyDeviceToHost, s));
        HIPCHECK(hipStreamSynchronize(s));
        float a = 0.f, b = 0.f;
        HIPCHECK(hipEventElapsedTime(&a, sim->ev[1], sim->ev[2]));
        HIPCHECK(hipEventElapsedTime(&b, sim->ev[2], sim->ev[3]));
        prop_ms += a; comp_ms += b;
        HIPCHECK(hipEventRecord(sim->ev[1], s));
        hits += nh;
    }
    if (taping) {
        mi_sim::TapeKey& tk = sim->tape_key;
        tk.valid = true; tk.event_id = event_id; tk.k = K; tk.photon_offset = photon_offset;
        tk.n_photon = n_photon; tk.n_gs = n_gs; tk.gs = d_gensteps; tk.generation = sim->generation;
    }
    sim->last_ms[1] = prop_ms;
    sim->last_ms[2] = comp_ms;
    float tot = 0.f;
    HIPCHECK(hipEventElapsedTime(&tot, sim->ev[0], sim->ev[3]));
    sim->last_ms[3] = tot;
    sim->last_ms[0] = std::max(0.0, (double)tot - prop_ms - comp_ms);
    if (n_hit_out) *n_hit_out = hits;
    if (d_hit && hits > hit_capacity) return fail(MI_ERR_CAPACITY, "hit_capacity too small: " + std::to_string(hits) + " hits");
    return MI_OK;
}

int mi_sim_simulate_device(mi_sim* sim, int32_t event_id, const void* d_gensteps, int64_t n_gs, int64_t n_photon,
                           int64_t photon_offset, void* d_photon, void* d_seq, void* d_hit, int64_t hit_capacity,
                           int64_t* n_hit_out, void* stream)
{
    return simulate_device_impl(sim, event_id, d_gensteps, n_gs, n_photon, photon_offset, d_photon, d_seq, d_hit,
                                hit_capacity, n_hit_out, nullptr, stream);
}

/*
 * No-allocation form of mi_sim_simulate_device: every buffer beyond the
 * caller's inputs and outputs lives in this GPU's static scratch array (its
 * address looked up at create), and the call makes no device or host
 * allocation. Differences from mi_sim_simulate_device: no max_slot slicing,
 * photons find their genstep by binary search of the prefix (same results),
 * limits of OA_NOALLOC_MAX_GS gensteps and OA_NOALLOC_MAX_PHOTON photons.
 */
int mi_sim_simulate_device_noalloc(mi_sim* sim, int32_t event_id, const void* d_gensteps, int64_t n_gs,
                                   int64_t n_photon, int64_t photon_offset, void* d_photon, void* d_seq, void* d_hit,
                                   int64_t hit_capacity, int64_t* n_hit_out, void* stream)
{
    if (!sim) return fail(MI_ERR_ARG, "null sim");
    if (n_hit_out) *n_hit_out = 0;
    if (n_gs <= 0) return fail(MI_NOTHING_TO_DO, "zero gensteps");
    if (!d_gensteps) return fail(MI_ERR_ARG, "null gensteps");
    if (d_hit && !d_photon) return fail(MI_ERR_ARG, "hits are gathered from the photon array: pass d_photon");
    if (n_gs > OA_NOALLOC_MAX_GS) return fail(MI_ERR_CAPACITY, "no-allocation simulate: more than 2^20 gensteps");
    if (n_photon > OA_NOALLOC_MAX_PHOTON) return fail(MI_ERR_CAPACITY, "no-allocation simulate: more than 2^26 photons");
    if (!sim->scratch) return fail(MI_ERR_HIP, "no-allocation simulate: no scratch address");
    hipStream_t s = stream ? (hipStream_t)stream : sim->stream;
    char* base = (char*)sim->scratch;
    int64_t* prefix = (int64_t*)(base + OA_NOALLOC_OFF_PREFIX);
    int64_t* ev = (int64_t*)(base + OA_NOALLOC_OFF_EV);
    int32_t* ev_id = (int32_t*)(base + OA_NOALLOC_OFF_EVID);
    unsigned long long* next = (unsigned long long*)(base + OA_NOALLOC_OFF_NEXT);
    uint32_t* xw = (uint32_t*)(base + OA_NOALLOC_OFF_XW);
    unsigned long long* words = (unsigned long long*)(base + OA_NOALLOC_OFF_WORDS);
    int32_t* word_prefix = (int32_t*)(base + OA_NOALLOC_OFF_WPREFIX);
    int64_t* block_total = (int64_t*)(base + OA_NOALLOC_OFF_BTOTAL);
    HIPCHECK(hipEventRecord(sim->ev[0], s));
    HIPCHECK(oa_launch::gs_prefix((const float*)d_gensteps, n_gs, prefix, s));
    int64_t total = 0;
    HIPCHECK(hipMemcpyAsync(&total, prefix + n_gs, sizeof(int64_t), hipMemcpyDeviceToHost, s));
    HIPCHECK(hipStreamSynchronize(s));
    if (n_photon > total) return fail(MI_ERR_ARG, "n_photon exceeds the gensteps' photons");
    if (n_photon < 0) n_photon = total;
    if (n_photon == 0) return fail(MI_NOTHING_TO_DO, "gensteps carry no photons");
    if (n_photon > OA_NOALLOC_MAX_PHOTON) return fail(MI_ERR_CAPACITY, "no-allocation simulate: more than 2^26 photons");
    int rc = check_precooked_rows(sim, photon_offset + n_photon);
    if (rc) return rc;
    const uint32_t* xw_table = nullptr;
    if (sim->cfg.rng_kind == MI_RNG_XORWOW) {
        const uint64_t skip = sim->cfg.rng_offset + sim->cfg.event_skipahead * (uint64_t)(uint32_t)event_id;
        oa_xorwow_jump_matrix(skip, sim->xw_host_J.data(), sim->xw_host_work.data());
        HIPCHECK(hipMemcpyAsync(xw, sim->xw_host_J.data(), sizeof(uint32_t) * OA_XORWOW_JUMP_WORDS,
                                hipMemcpyHostToDevice, s));
        xw_table = xw;
    }
    HIPCHECK(oa_launch::noalloc_setup(ev, ev_id, next, n_gs, n_photon, event_id, s));
    HIPCHECK(hipEventRecord(sim->ev[1], s));
    const int64_t n_words = (n_photon + 63) / 64;
    SimParams P = base_params(sim);
    P.xw_jump = xw_table;
    P.gensteps = (const float*)d_gensteps;
    P.gs_prefix = prefix;
    P.seed_map = nullptr;                 /* genstep of each photon by binary search of the prefix */
    P.gs_event = nullptr;
    P.n_gs = n_gs;
    P.n_event = 1;
    P.ev_gs_begin = ev;
    P.ev_ph_begin = ev + 2;
    P.ev_id = ev_id;
    P.n_photon = n_photon;
    P.photon_offset = photon_offset;
    P.photon = (float4*)d_photon;
    P.seq = (uint64_t*)d_seq;
    P.photonlite = nullptr;
    P.hitwords = words;
    P.next = next;
    if (sim->cfg.kernel == MI_KERNEL_WAVEFRONT) HIPCHECK(oa_launch::zero_u64(words, n_words, s));
    HIPCHECK(oa_launch::simulate(P, sim->cfg.kernel, sim->cfg.block_size, wave_grid(sim), s));
    HIPCHECK(hipEventRecord(sim->ev[2], s));
    const int64_t nb = (n_words + 255) / 256;
    HIPCHECK(oa_launch::compact_hits(words, n_words, word_prefix, block_total, (const float4*)d_photon,
                                     (float4*)d_hit, d_hit ? hit_capacity : 0, s));
    HIPCHECK(hipEventRecord(sim->ev[3], s));
    int64_t hits = 0;
    HIPCHECK(hipMemcpyAsync(&hits, block_total + nb, sizeof(int64_t), hipMemcpyDeviceToHost, s));
    HIPCHECK(hipStreamSynchronize(s));
    float a = 0.f, b = 0.f, c = 0.f, t = 0.f;
    HIPCHECK(hipEventElapsedTime(&a, sim->ev[0], sim->ev[1]));
    HIPCHECK(hipEventElapsedTime(&b, sim->ev[1], sim->ev[2]));
    HIPCHECK(hipEventElapsedTime(&c, sim->ev[2], sim->ev[3]));
    HIPCHECK(hipEventElapsedTime(&t, sim->ev[0], sim->ev[3]));
    sim->last_ms[0] = a; sim->last_ms[1] = b; sim->last_ms[2] = c; sim->last_ms[3] = t;
    if (n_hit_out) *n_hit_out = hits;
    if (d_hit && hits > hit_capacity) return fail(MI_ERR_CAPACITY, "hit_capacity too small");
    return MI_OK;
}

int mi_sim_simulate_lite_device(mi_sim* sim, int32_t event_id, const void* d_gensteps, int64_t n_gs, int64_t n_photon,
                                int64_t photon_offset, void* d_photonlite, void* d_photon, void* d_seq, void* d_hit,
                                int64_t hit_capacity, int64_t* n_hit_out, void* stream)
{
    if (!d_photonlite) return fail(MI_ERR_ARG, "null photonlite buffer");
    return simulate_device_impl(sim, event_id, d_gensteps, n_gs, n_photon, photon_offset, d_photon, d_seq, d_hit,
                                hit_capacity, n_hit_out, d_photonlite, stream);
}

int mi_sim_simulate_batch_device(mi_sim* sim, int32_t n_event, const int32_t* event_ids, const int64_t* gs_offsets,
                                 const void* d_gensteps, int64_t* photon_offsets, void* d_photon, void* d_seq,
                                 void* d_hit, int64_t hit_capacity, int64_t* hit_offsets, void* stream)
{
    if (!sim || n_event <= 0 || !event_ids || !gs_offsets || !photon_offsets || !hit_offsets)
        return fail(MI_ERR_ARG, "bad batch arguments");
    if (d_hit && !d_photon) return fail(MI_ERR_ARG, "hits are gathered from the photon array: pass d_photon");
    const int64_t n_gs = gs_offsets[n_event];
    if (n_gs <= 0) return fail(MI_NOTHING_TO_DO, "zero gensteps");
    hipStream_t s = stream ? (hipStream_t)stream : sim->stream;
    HIPCHECK(hipEventRecord(sim->ev[0], s));
    HIPCHECK(sim->prefix.reserve((size_t)n_gs + 1));
    HIPCHECK(oa_launch::gs_prefix((const float*)d_gensteps, n_gs, sim->prefix.p, s));
    std::vector<int64_t> hp((size_t)n_gs + 1);
    HIPCHECK(hipMemcpyAsync(hp.data(), sim->prefix.p, sizeof(int64_t) * (n_gs + 1), hipMemcpyDeviceToHost, s));
    HIPCHECK(hipStreamSynchronize(s));
    std::vector<int64_t> evh(2 * ((size_t)n_event + 1));
    for (int e = 0; e <= n_event; ++e) {
        evh[e] = gs_offsets[e];
        evh[n_event + 1 + e] = hp[gs_offsets[e]];
        photon_offsets[e] = hp[gs_offsets[e]];
    }
    const int64_t n_photon = hp[n_gs];
    if (n_photon == 0) return fail(MI_NOTHING_TO_DO, "gensteps carry no photons");
    int64_t max_ev = 0;
    for (int e = 0; e < n_event; ++e) max_ev = std::max(max_ev, photon_offsets[e + 1] - photon_offsets[e]);
    int prc = check_precooked_rows(sim, max_ev);
    if (prc) return prc;
    const uint32_t* xw_table = nullptr;
    prc = xorwow_table(sim, event_ids, n_event, sim->xw_jump, &xw_table);
    if (prc) return prc;
    HIPCHECK(sim->evbuf.reserve(evh.size() + n_event + 1));
    HIPCHECK(sim->ev_id.reserve((size_t)n_event));
    HIPCHECK(hipMemcpyAsync(sim->evbuf.p, evh.data(), sizeof(int64_t) * evh.size(), hipMemcpyHostToDevice, s));
    HIPCHECK(hipMemcpyAsync(sim->ev_id.p, event_ids, sizeof(int32_t) * n_event, hipMemcpyHostToDevice, s));
    HIPCHECK(hipEventRecord(sim->ev[1], s));
    SimParams P = base_params(sim);
    P.xw_jump = xw_table;
    P.gensteps = (const float*)d_gensteps;
    P.gs_prefix = sim->prefix.p;
    P.n_gs = n_gs;
    P.n_event = n_event;
    P.ev_gs_begin = sim->evbuf.p;
    P.ev_ph_begin = sim->evbuf.p + (n_event + 1);
    P.ev_id = sim->ev_id.p;
    P.n_photon = n_photon;
    P.photon_offset = 0;
    P.photon = (float4*)d_photon;
    P.seq = (uint64_t*)d_seq;
    int rc = launch(sim, P, s);
    if (rc) return rc;
    HIPCHECK(hipEventRecord(sim->ev[2], s));
    const int64_t n_words = (n_photon + 63) / 64;
    const int64_t nb = (n_words + 255) / 256;
    HIPCHECK(sim->word_prefix.reserve((size_t)n_words + 1));
    HIPCHECK(sim->block_total.reserve((size_t)nb + 1));
    HIPCHECK(oa_launch::compact_hits(sim->words.p, n_words, sim->word_prefix.p, sim->block_total.p,
                                     (const float4*)d_photon, (float4*)d_hit, d_hit ? hit_capacity : 0, s));
    DBuf<int64_t>& hob = sim->evbuf;   /* reuse tail of evbuf for the offsets */
    int64_t* d_ho = hob.p + evh.size();
    HIPCHECK(oa_launch::event_hit_offsets(sim->words.p, sim->word_prefix.p, sim->block_total.p,
                                          P.ev_ph_begin, n_event, n_words, d_ho, s));
    HIPCHECK(hipMemcpyAsync(hit_offsets, d_ho, sizeof(int64_t) * (n_event + 1), hipMemcpyDeviceToHost, s));
    HIPCHECK(hipEventRecord(sim->ev[3], s));
    HIPCHECK(hipEventSynchronize(sim->ev[3]));
    float a = 0.f, b = 0.f, c = 0.f, t = 0.f;
    HIPCHECK(hipEventElapsedTime(&a, sim->ev[0], sim->ev[1]));
    HIPCHECK(hipEventElapsedTime(&b, sim->ev[1], sim->ev[2]));
    HIPCHECK(hipEventElapsedTime(&c, sim->ev[2], sim->ev[3]));
    HIPCHECK(hipEventElapsedTime(&t, sim->ev[0], sim->ev[3]));
    sim->last_ms[0] = a; sim->last_ms[1] = b; sim->last_ms[2] = c; sim->last_ms[3] = t;
    if (d_hit && hit_offsets[n_event] > hit_capacity)
        return fail(MI_ERR_CAPACITY, "hit_capacity too small: " + std::to_string(hit_offsets[n_event]) + " hits");
    return MI_OK;
}

int mi_sim_simulate(mi_sim* sim, int32_t event_id, const float* gensteps, int64_t n_gs, mi_event_out* out)
{
    if (!sim || !out) return fail(MI_ERR_ARG, "null");
    std::memset(out, 0, sizeof(*out));
    if (n_gs <= 0) return fail(MI_NOTHING_TO_DO, "zero gensteps");
    if (!gensteps) return fail(MI_ERR_ARG, "null gensteps");
    int64_t n_photon = 0;
    for (int64_t g = 0; g < n_gs; ++g) {
        uint32_t np;
        std::memcpy(&np, gensteps + 24 * g + 3, 4);
        n_photon += np;
    }
    if (n_photon == 0) return fail(MI_NOTHING_TO_DO, "gensteps carry no photons");
    hipStream_t s = sim->stream;
    HIPCHECK(upload(sim->gs, gensteps, (size_t)n_gs * 24, s));
    const uint32_t om = sim->cfg.output_mask;
    HIPCHECK(sim->photon.reserve((size_t)n_photon * 4));
    if (om & MI_OUT_SEQ) HIPCHECK(sim->seq.reserve((size_t)n_photon * 4));
    if (om & MI_OUT_HIT) HIPCHECK(sim->hit.reserve((size_t)n_photon * 4));
    int64_t n_hit = 0;
    int rc = mi_sim_simulate_device(sim, event_id, sim->gs.p, n_gs, n_photon, 0, sim->photon.p,
                                    (om & MI_OUT_SEQ) ? sim->seq.p : nullptr,
                                    (om & MI_OUT_HIT) ? sim->hit.p : nullptr, n_photon, &n_hit, s);
    if (rc) return rc;
    out->n_photon = n_photon;
    out->n_hit = n_hit;
    out->device_ms = sim->last_ms[3];
    if (om & MI_OUT_PHOTON) {
        sim->h_photon.resize((size_t)n_photon * 16);
        HIPCHECK(hipMemcpy(sim->h_photon.data(), sim->photon.p, sizeof(float) * 16 * n_photon, hipMemcpyDeviceToHost));
        out->photon = sim->h_photon.data();
    }
    if (om & MI_OUT_SEQ) {
        sim->h_seq.resize((size_t)n_photon * 4);
        HIPCHECK(hipMemcpy(sim->h_seq.data(), sim->seq.p, sizeof(uint64_t) * 4 * n_photon, hipMemcpyDeviceToHost));
        out->seq = sim->h_seq.data();
    }
    if (om & MI_OUT_HIT) {
        sim->h_hit.resize((size_t)std::max<int64_t>(n_hit, 1) * 16);
        if (n_hit) HIPCHECK(hipMemcpy(sim->h_hit.data(), sim->hit.p, sizeof(float) * 16 * n_hit, hipMemcpyDeviceToHost));
        out->hit = sim->h_hit.data();
    }
    return MI_OK;
}

int mi_sim_reset(mi_sim* sim, int32_t event_id)
{
    (void)event_id;
    if (!sim) return fail(MI_ERR_ARG, "null sim");
    std::vector<float>().swap(sim->h_photon);
    std::vector<float>().swap(sim->h_hit);
    std::vector<float>().swap(sim->h_simtrace);
    std::vector<uint64_t>().swap(sim->h_seq);
    return MI_OK;
}

int mi_sim_simtrace_device(mi_sim* sim, const void* d_gensteps, int64_t n_gs, int64_t n_photon, void* d_simtrace,
                           void* stream)
{
    if (!sim) return fail(MI_ERR_ARG, "null sim");
    if (n_gs <= 0) return fail(MI_NOTHING_TO_DO, "zero gensteps");
    if (!d_gensteps || !d_simtrace) return fail(MI_ERR_ARG, "null simtrace buffers");
    hipStream_t s = stream ? (hipStream_t)stream : sim->stream;
    HIPCHECK(sim->prefix.reserve((size_t)n_gs + 1));
    HIPCHECK(oa_launch::gs_prefix((const float*)d_gensteps, n_gs, sim->prefix.p, s));
    int64_t total = 0;
    HIPCHECK(hipMemcpyAsync(&total, sim->prefix.p + n_gs, sizeof(int64_t), hipMemcpyDeviceToHost, s));
    HIPCHECK(hipStreamSynchronize(s));
    if (n_photon >= 0 && n_photon != total) return fail(MI_ERR_ARG, "n_photon does not match the gensteps");
    if (total == 0) return fail(MI_NOTHING_TO_DO, "gensteps carry no photons");
    SimParams P = base_params(sim);
    P.gensteps = (const float*)d_gensteps;
    P.gs_prefix = sim->prefix.p;
    P.n_gs = n_gs;
    P.n_photon = total;
    HIPCHECK(oa_launch::simtrace(P, (float4*)d_simtrace, s));
    HIPCHECK(hipStreamSynchronize(s));
    return MI_OK;
}

int mi_sim_simtrace(mi_sim* sim, const float* gensteps, int64_t n_gs, const float** simtrace, int64_t* n_simtrace,
                    double* device_ms)
{
    if (!sim || !simtrace || !n_simtrace) return fail(MI_ERR_ARG, "null");
    *simtrace = nullptr;
    *n_simtrace = 0;
    if (device_ms) *device_ms = 0.0;
    if (n_gs <= 0) return fail(MI_NOTHING_TO_DO, "zero gensteps");
    if (!gensteps) return fail(MI_ERR_ARG, "null gensteps");
    int64_t n = 0;
    for (int64_t g = 0; g < n_gs; ++g) {
        uint32_t np;
        std::memcpy(&np, gensteps + 24 * g + 3, 4);
        n += np;
    }
    if (n == 0) return fail(MI_NOTHING_TO_DO, "gensteps carry no photons");
    hipStream_t s = sim->stream;
    HIPCHECK(upload(sim->gs, gensteps, (size_t)n_gs * 24, s));
    HIPCHECK(sim->photon.reserve((size_t)n * 4));
    HIPCHECK(hipEventRecord(sim->ev[0], s));
    int rc = mi_sim_simtrace_device(sim, sim->gs.p, n_gs, n, sim->photon.p, s);
    if (rc) return rc;
    HIPCHECK(hipEventRecord(sim->ev[3], s));
    HIPCHECK(hipEventSynchronize(sim->ev[3]));
    float ms = 0.f;
    HIPCHECK(hipEventElapsedTime(&ms, sim->ev[0], sim->ev[3]));
    sim->h_simtrace.resize((size_t)n * 16);
    HIPCHECK(hipMemcpy(sim->h_simtrace.data(), sim->photon.p, sizeof(float) * 16 * n, hipMemcpyDeviceToHost));
    *simtrace = sim->h_simtrace.data();
    *n_simtrace = n;
    if (device_ms) *device_ms = ms;
    return MI_OK;
}

int mi_sim_trace_device(mi_sim* sim, const void* d_rays, int64_t n, void* d_isect, void* stream)
{
    if (!sim || !d_rays || !d_isect || n < 0) return fail(MI_ERR_ARG, "bad trace arguments");
    if (n == 0) return MI_OK;
    hipStream_t s = stream ? (hipStream_t)stream : sim->stream;
    SimParams P = base_params(sim);
    HIPCHECK(oa_launch::trace(P, (const float4*)d_rays, n, (float4*)d_isect, s));
    HIPCHECK(hipStreamSynchronize(s));
    return MI_OK;
}

int mi_sim_refit_instances(mi_sim* sim, const float* inst, int64_t n_inst)
{
    if (!sim || !inst) return fail(MI_ERR_ARG, "null");
    if (n_inst != sim->n_inst) return fail(MI_ERR_ARG, "refit keeps the instance count");
    hipStream_t s = sim->stream;
    std::vector<InstRec> rec(n_inst);
    std::vector<InstRec> old(n_inst);
    HIPCHECK(hipMemcpy(old.data(), sim->inst.p, sizeof(InstRec) * n_inst, hipMemcpyDeviceToHost));
    for (int64_t i = 0; i < n_inst; ++i) {
        const float* m = inst + 16 * i;
        int32_t gas;
        std::memcpy(&gas, m + 7, 4);
        if (gas < 0 || (size_t)gas >= sim->solid_blas_root.size()) return fail(MI_ERR_GEOMETRY, "gas index");
        if (sim->solid_blas_root[gas] != sim->h_inst_blas_root[i]) return fail(MI_ERR_ARG, "refit keeps each instance's solid");
        invert_affine16(m, rec[i].m);
        rec[i].m[3] = old[i].m[3];
        int32_t ident;
        std::memcpy(&ident, m + 11, 4);
        rec[i].m[7] = f32_of((uint32_t)ident);
        rec[i].m[11] = old[i].m[11];
    }
    HIPCHECK(upload(sim->inst, rec.data(), rec.size(), s));
    sim->h_o2w.assign(inst, inst + 16 * n_inst);
    /* the instances' world boxes as a host build makes them (exact leaf bounds), rounded
       outward to float; the TLAS node boxes are then refitted level by level on the GPU */
    std::vector<oa_bvh::Item> items;
    instance_items(sim, items);
    std::vector<float> boxes((size_t)n_inst * 6);
    for (int64_t i = 0; i < n_inst; ++i)
        oa_bvh::Builder::store(boxes.data() + 6 * i, boxes.data() + 6 * i + 3, items[i].lo, items[i].hi);
    HIPCHECK(upload(sim->inst_box, boxes.data(), boxes.size(), s));
    HIPCHECK(oa_launch::refit(sim->tlas.p, sim->n_tlas, sim->tlas_depth, sim->tlas_items.p, sim->inst_box.p, s));
    /* the free ball of the moved boxes, written in place (captured batches read it there) */
    const float4 ball = free_ball(sim, items, sim->h_is_front);
    HIPCHECK(hipMemcpyAsync(sim->free_ball.p, &ball, sizeof(ball), hipMemcpyHostToDevice, s));
    HIPCHECK(hipStreamSynchronize(s));
    return MI_OK;
}

/* score-row scratch limit of one adjoint chunk, in floats (4 GiB) */
static const int64_t OA_ADJ_ROWS_MAX = (int64_t)1 << 30;

int mi_sim_set_bnd(mi_sim* sim, const float* bnd)
{
    if (!sim || !bnd) return fail(MI_ERR_ARG, "null sim or bnd");
    const size_t nb = (size_t)sim->n_bnd * 4 * 2 * sim->n_wl;
    HIPCHECK(hipMemcpyAsync(sim->bnd.p, bnd, nb * sizeof(float4), hipMemcpyHostToDevice, sim->stream));
    HIPCHECK(hipStreamSynchronize(sim->stream));
    return MI_OK;
}

static int check_adjoint(const mi_sim* sim, const mi_adjoint* adj)
{
    if (adj->n_bin <= 0 || adj->n_block <= 0 || adj->n_det <= 0 || adj->n_det > 0xffff || adj->n_mat < 0)
        return fail(MI_ERR_ARG, "adjoint: n_bin, n_block must be > 0, n_det in [1, 65535], n_mat >= 0");
    if (!adj->mat_block || !adj->sample_bin) return fail(MI_ERR_ARG, "adjoint: null mat_block or sample_bin");
    for (int32_t m = 0; m <= adj->n_mat; ++m)
        if (adj->mat_block[m] < -1 || adj->mat_block[m] >= adj->n_block)
            return fail(MI_ERR_ARG, "adjoint: mat_block entry out of [-1, n_block)");
    for (int32_t i = 0; i < sim->n_wl; ++i)
        if (adj->sample_bin[i] < 0 || adj->sample_bin[i] >= adj->n_bin)
            return fail(MI_ERR_ARG, "adjoint: sample_bin entry out of [0, n_bin)");
    const int64_t n_param64 = (int64_t)adj->n_block * 2 * adj->n_bin;
    if (n_param64 > (1 << 20)) return fail(MI_ERR_ARG, "adjoint: too many parameters");
    return MI_OK;
}

int mi_sim_adjoint_device(mi_sim* sim, int32_t event_id, const void* d_gensteps, int64_t n_gs, int64_t n_photon,
                          int64_t photon_offset, const void* d_hit, int64_t n_hit, const mi_adjoint* adj,
                          void* d_grad, void* stream)
{
    if (!sim || !adj) return fail(MI_ERR_ARG, "null sim or adjoint description");
    if (!d_grad) return fail(MI_ERR_ARG, "null gradient buffer");
    if (n_gs <= 0 || !d_gensteps) return fail(MI_ERR_ARG, "the event's gensteps are required");
    if (n_hit < 0 || (n_hit > 0 && !d_hit)) return fail(MI_ERR_ARG, "bad hit array");
    if (n_hit > (int64_t)INT32_MAX) return fail(MI_ERR_CAPACITY, "more than 2^31-1 hits in one adjoint call");
    int arc = check_adjoint(sim, adj);
    if (arc) return arc;
    const int32_t n_param = adj->n_block * 2 * adj->n_bin;
    hipStream_t s = stream ? (hipStream_t)stream : sim->stream;
    HIPCHECK(upload(sim->adj_mat_block, adj->mat_block, (size_t)adj->n_mat + 1, s));
    HIPCHECK(upload(sim->adj_sample_bin, adj->sample_bin, (size_t)sim->n_wl, s));
    HIPCHECK(sim->prefix.reserve((size_t)n_gs + 1));
    HIPCHECK(oa_launch::gs_prefix((const float*)d_gensteps, n_gs, sim->prefix.p, s));
    if (n_photon < 0) {
        HIPCHECK(hipMemcpyAsync(&n_photon, sim->prefix.p + n_gs, sizeof(int64_t), hipMemcpyDeviceToHost, s));
        HIPCHECK(hipStreamSynchronize(s));
    }
    int prc = check_precooked_rows(sim, photon_offset + n_photon);
    if (prc) return prc;
    const uint32_t* xw_table = nullptr;
    prc = xorwow_table(sim, &event_id, 1, sim->xw_jump, &xw_table);
    if (prc) return prc;
    int64_t evh[4] = { 0, n_gs, 0, n_photon };
    int32_t evid = event_id;
    HIPCHECK(sim->evbuf.reserve(4));
    HIPCHECK(sim->ev_id.reserve(1));
    HIPCHECK(hipMemcpyAsync(sim->evbuf.p, evh, sizeof(evh), hipMemcpyHostToDevice, s));
    HIPCHECK(hipMemcpyAsync(sim->ev_id.p, &evid, sizeof(evid), hipMemcpyHostToDevice, s));
    SimParams P = base_params(sim);
    P.xw_jump = xw_table;
    P.gensteps = (const float*)d_gensteps;
    P.gs_prefix = sim->prefix.p;
    P.n_gs = n_gs;
    P.n_event = 1;
    P.ev_gs_begin = sim->evbuf.p;
    P.ev_ph_begin = sim->evbuf.p + 2;
    P.ev_id = sim->ev_id.p;
    P.n_photon = n_photon;
    P.photon_offset = photon_offset;
    const mi_sim::TapeKey& tk = sim->tape_key;
    if (tk.valid && tk.event_id == event_id && tk.photon_offset == photon_offset && tk.n_photon == n_photon &&
        tk.n_gs == n_gs && tk.gs == d_gensteps && tk.generation == sim->generation &&
        oa_launch::tape_reduce_rows(n_param) > 0) {
        HIPCHECK(sim->tape_flag.reserve(1));
        P.tape = sim->tape.p; P.tape_n = sim->tape_n.p; P.tape_stride = tk.n_photon; P.tape_k = tk.k;
        P.tape_overflow = sim->tape_flag.p;
    }
    const uint32_t n_det = (uint32_t)adj->n_det;
    /* hits are processed in chunks of consecutive hits (photon order) whose
       score rows fit OA_ADJ_ROWS_MAX floats; the per-detector Kahan sums
       carry their compensation from chunk to chunk */
    /* with a tape of at least max_bounce records no photon overflows it (one segment per
       step): no score rows are needed, so the hits are not chunked unless the caller asks */
    const bool tape_only = P.tape && P.tape_k >= sim->cfg.max_bounce;
    const int64_t rows_max = adj->max_row_floats > 0 ? adj->max_row_floats : OA_ADJ_ROWS_MAX;
    const int64_t chunk = (tape_only && adj->max_row_floats <= 0)
                              ? std::max<int64_t>(n_hit, 1)
                              : std::max<int64_t>(1, std::min<int64_t>(std::max<int64_t>(n_hit, 1), rows_max / n_param));
    const size_t tmp_bytes = std::max<size_t>(1, oa_launch::adjoint_sort_bytes(chunk, n_det));
    if (sim->adj_rows.reserve((size_t)((tape_only ? 1 : chunk) * n_param)) != hipSuccess)
        return fail(MI_ERR_CAPACITY, "adjoint: cannot allocate " + std::to_string(chunk * n_param * 4) + " bytes of score rows");
    HIPCHECK(sim->adj_comp.reserve((size_t)n_det * n_param));
    HIPCHECK(sim->adj_key.reserve((size_t)chunk));
    HIPCHECK(sim->adj_val.reserve((size_t)chunk));
    HIPCHECK(sim->adj_key2.reserve((size_t)chunk));
    HIPCHECK(sim->adj_val2.reserve((size_t)chunk));
    HIPCHECK(sim->adj_tmp.reserve(tmp_bytes));
    HIPCHECK(sim->adj_seg.reserve((size_t)n_det + 1));
    /* persistent replay over every resident block slot */
    const int replay_grid = sim->n_cu * std::max(1, oa_launch::adjoint_blocks_per_cu(sim->cfg.rng_kind));
    double replay_ms = 0.0, reduce_ms = 0.0;
    oa_launch::AdjBufs ab;
    ab.mat_block = sim->adj_mat_block.p; ab.n_mat = adj->n_mat;
    ab.sample_bin = sim->adj_sample_bin.p; ab.n_bin = adj->n_bin; ab.n_param = n_param;
    ab.hit = (const float4*)d_hit; ab.hit_off = nullptr; ab.n_ev = 0; ab.hit_cap = 0;
    ab.n_det = n_det; ab.S = sim->adj_rows.p;
    ab.key = sim->adj_key.p; ab.val = sim->adj_val.p; ab.key2 = sim->adj_key2.p; ab.val2 = sim->adj_val2.p;
    ab.tmp = sim->adj_tmp.p; ab.tmp_bytes = tmp_bytes; ab.seg = sim->adj_seg.p;
    ab.G = (float*)d_grad; ab.C = sim->adj_comp.p;
    int64_t h0 = 0;
    do {
        const int64_t nh = std::min<int64_t>(chunk, n_hit - h0);
        HIPCHECK(hipEventRecord(sim->ev[0], s));
        HIPCHECK(oa_launch::adjoint_chunk(P, ab, h0, nh, h0 == 0 ? 1 : 0, replay_grid, sim->ev[1], s));
        HIPCHECK(hipEventRecord(sim->ev[2], s));
        HIPCHECK(hipEventSynchronize(sim->ev[2]));
        float a = 0.f, b = 0.f;
        HIPCHECK(hipEventElapsedTime(&a, sim->ev[0], sim->ev[1]));
        HIPCHECK(hipEventElapsedTime(&b, sim->ev[1], sim->ev[2]));
        replay_ms += a; reduce_ms += b;
        h0 += nh;
    } while (h0 < n_hit);
    sim->adj_ms[0] = replay_ms; sim->adj_ms[1] = reduce_ms; sim->adj_ms[2] = replay_ms + reduce_ms;
    return MI_OK;
}

int mi_sim_merge_hits_device(mi_sim* sim, int32_t record, const void* d_in, int64_t n_in, uint32_t select_mask,
                             float time_window, void* d_out, int64_t out_capacity, int64_t* n_out, void* stream)
{
    if (!sim || !n_out) return fail(MI_ERR_ARG, "null argument");
    *n_out = 0;
    if (record != MI_RECORD_PHOTON && record != MI_RECORD_PHOTONLITE) return fail(MI_ERR_ARG, "unknown record type");
    if (n_in < 0 || (n_in > 0 && !d_in)) return fail(MI_ERR_ARG, "bad input records");
    if (n_in > (int64_t)INT32_MAX) return fail(MI_ERR_CAPACITY, "more than 2^31-1 records in one merge");
    if (!(time_window >= 0.f)) return fail(MI_ERR_ARG, "time_window must be >= 0");
    if (out_capacity < 0 || (out_capacity > 0 && !d_out)) return fail(MI_ERR_ARG, "bad output buffer");
    if (n_in == 0) return MI_OK;
    hipStream_t s = stream ? (hipStream_t)stream : sim->stream;
    const size_t n = (size_t)n_in;
    HIPCHECK(sim->m_flags.reserve(n));
    HIPCHECK(sim->m_sel.reserve(n));
    HIPCHECK(sim->m_vals.reserve(n));
    HIPCHECK(sim->m_vals2.reserve(n));
    HIPCHECK(sim->m_keys.reserve(n));
    HIPCHECK(sim->m_keys2.reserve(n));
    HIPCHECK(sim->m_len.reserve(n));
    HIPCHECK(sim->m_off.reserve(n));
    HIPCHECK(sim->m_counts.reserve(2));
    const size_t tmp_bytes = std::max<size_t>(1, oa_merge::scratch_bytes(n_in));
    HIPCHECK(sim->m_tmp.reserve(tmp_bytes));
    oa_merge::Scratch S{ sim->m_flags.p, sim->m_sel.p, sim->m_vals.p, sim->m_vals2.p, sim->m_keys.p, sim->m_keys2.p,
                         sim->m_len.p, sim->m_off.p, sim->m_counts.p, sim->m_tmp.p, tmp_bytes };
    const int rec = record == MI_RECORD_PHOTON ? OA_MERGE_PHOTON : OA_MERGE_PHOTONLITE;
    HIPCHECK(oa_merge::select(S, rec, d_in, n_in, select_mask, s));
    int32_t cnt[2] = { 0, 0 };
    HIPCHECK(hipMemcpyAsync(cnt, sim->m_counts.p, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    HIPCHECK(hipStreamSynchronize(s));
    const int64_t n_sel = cnt[0];
    if (n_sel == 0) return MI_OK;
    if (time_window == 0.f) {
        /* selection only, input order (SPM::NOMERGE_TIME_WINDOW) */
        *n_out = n_sel;
        if (n_sel > out_capacity) return fail(MI_ERR_CAPACITY, "out_capacity too small: " + std::to_string(n_sel) + " records");
        HIPCHECK(oa_merge::gather_selected(S, rec, d_in, d_out, n_sel, s));
        HIPCHECK(hipStreamSynchronize(s));
        return MI_OK;
    }
    HIPCHECK(oa_merge::sort_runs(S, rec, d_in, n_sel, time_window, s));
    HIPCHECK(hipMemcpyAsync(cnt + 1, sim->m_counts.p + 1, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    HIPCHECK(hipStreamSynchronize(s));
    const int64_t n_runs = cnt[1];
    *n_out = n_runs;
    if (n_runs > out_capacity) return fail(MI_ERR_CAPACITY, "out_capacity too small: " + std::to_string(n_runs) + " records");
    HIPCHECK(oa_merge::reduce_runs(S, rec, d_in, d_out, n_runs, s));
    HIPCHECK(hipStreamSynchronize(s));
    return MI_OK;
}

int mi_sim_last_adjoint_timing(const mi_sim* sim, double* ms3)
{
    if (!sim || !ms3) return fail(MI_ERR_ARG, "null argument");
    for (int i = 0; i < 3; ++i) ms3[i] = sim->adj_ms[i];
    return MI_OK;
}

/* ------------------------------------------------------ captured batches */

struct mi_batch {
    mi_sim* sim = nullptr;
    uint64_t generation = 0;
    int32_t n_event = 0;
    int64_t n_gs = 0, n_photon = 0, n_words = 0, hit_capacity = 0;
    bool want_hits = false;
    hipStream_t cap = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hipEvent_t ev[2] = { nullptr, nullptr };
    DBuf<int64_t> prefix, saved_prefix, evbuf, block_total, ho;
    DBuf<int32_t> ev_id, word_prefix, gs_event;
    DBuf<uint32_t> seed, xw_jump;
    DBuf<unsigned long long> words, next, flag;
    int64_t* h_out = nullptr;   /* pinned: hit offsets (n_event+1), then the prefix-mismatch flag */
    bool launched = false;
    double last_ms = 0.0;
    /* adjoint part (optional): replay + per-detector reduction of every hit, chunked */
    bool has_adj = false;
    int32_t n_mat = 0, n_bin = 0, n_param = 0, n_det = 0;
    int64_t chunk = 0, n_chunk = 0;
    size_t tmp_bytes = 0;
    float* d_grad = nullptr;
    DBuf<int32_t> mat_block, sample_bin;
    DBuf<float> rows, comp;
    DBuf<uint32_t> key, val, key2, val2;
    DBuf<unsigned char> tmp;
    DBuf<int64_t> seg;
    DBuf<uint32_t> tape, tape_flag;     /* segment records of the forward (mi_config.adjoint_tape) */
    DBuf<uint8_t> tape_n;
    hipEvent_t ev_mid = nullptr;
    double last_fwd_ms = 0.0, last_adj_ms = 0.0;
};

int mi_batch_destroy(mi_batch* b)
{
    if (!b) return MI_OK;
    if (b->launched && b->ev[1]) (void)hipEventSynchronize(b->ev[1]);
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    for (auto& e : b->ev) if (e) (void)hipEventDestroy(e);
    if (b->ev_mid) (void)hipEventDestroy(b->ev_mid);
    if (b->cap) (void)hipStreamDestroy(b->cap);
    if (b->h_out) (void)hipHostFree(b->h_out);
    b->prefix.release(); b->saved_prefix.release(); b->evbuf.release(); b->block_total.release(); b->ho.release();
    b->ev_id.release(); b->word_prefix.release(); b->words.release(); b->next.release(); b->flag.release();
    b->seed.release(); b->gs_event.release(); b->xw_jump.release();
    b->mat_block.release(); b->sample_bin.release(); b->rows.release(); b->comp.release();
    b->key.release(); b->val.release(); b->key2.release(); b->val2.release(); b->tmp.release(); b->seg.release();
    b->tape.release(); b->tape_n.release(); b->tape_flag.release();
    delete b;
    return MI_OK;
}

/* the work of one batch launch, issued on stream s (recorded when s is capturing) */
static hipError_t batch_work(mi_batch* b, const SimParams& P, const void* d_gensteps, const void* d_photon, void* d_hit,
                             hipStream_t s)
{
    const mi_sim* sim = b->sim;
    hipError_t e;
    if ((e = oa_launch::gs_prefix((const float*)d_gensteps, b->n_gs, b->prefix.p, s)) != hipSuccess) return e;
    if ((e = oa_launch::zero_u64(b->flag.p, 1, s)) != hipSuccess) return e;
    if ((e = oa_launch::prefix_check(b->prefix.p, b->saved_prefix.p, b->n_gs + 1, b->flag.p, s)) != hipSuccess) return e;
    /* the seed map follows the captured prefix; a changed prefix is reported by the flag and the
       replay fails on the host, so its photons are never used */
    if ((e = oa_launch::seed_map(b->saved_prefix.p, b->n_gs, b->n_photon, b->seed.p, P.ev_gs_begin, b->n_event,
                                 b->gs_event.p, s)) != hipSuccess) return e;
    if (sim->cfg.kernel == MI_KERNEL_WAVEFRONT) {
        if ((e = oa_launch::zero_u64(b->words.p, b->n_words, s)) != hipSuccess) return e;
        if ((e = oa_launch::zero_u64(b->next.p, 1, s)) != hipSuccess) return e;
    }
    if ((e = oa_launch::simulate(P, sim->cfg.kernel, sim->cfg.block_size, wave_grid(sim), s)) != hipSuccess) return e;
    if ((e = oa_launch::compact_hits(b->words.p, b->n_words, b->word_prefix.p, b->block_total.p,
                                     (const float4*)d_photon, (float4*)d_hit, d_hit ? b->hit_capacity : 0, s)) != hipSuccess)
        return e;
    if ((e = oa_launch::event_hit_offsets(b->words.p, b->word_prefix.p, b->block_total.p, P.ev_ph_begin, b->n_event,
                                          b->n_words, b->ho.p, s)) != hipSuccess)
        return e;
    if ((e = hipMemcpyAsync(b->h_out, b->ho.p, sizeof(int64_t) * (b->n_event + 1), hipMemcpyDeviceToHost, s)) != hipSuccess)
        return e;
    return hipMemcpyAsync(b->h_out + b->n_event + 1, b->flag.p, sizeof(int64_t), hipMemcpyDeviceToHost, s);
}

/* adjoint of every hit of the batch: fixed chunks of b->chunk hit slots; the
   kernels read the hit count from the device offsets, so no host round trip */
static hipError_t batch_adjoint_work(mi_batch* b, SimParams P, const void* d_hit, hipStream_t s)
{
    const mi_sim* sim = b->sim;
    const int grid = sim->n_cu * std::max(1, oa_launch::adjoint_blocks_per_cu(sim->cfg.rng_kind));
    hipError_t e;
    oa_launch::AdjBufs ab;
    ab.mat_block = b->mat_block.p; ab.n_mat = b->n_mat;
    ab.sample_bin = b->sample_bin.p; ab.n_bin = b->n_bin; ab.n_param = b->n_param;
    ab.hit = (const float4*)d_hit; ab.hit_off = b->ho.p; ab.n_ev = b->n_event; ab.hit_cap = b->hit_capacity;
    ab.n_det = (uint32_t)b->n_det; ab.S = b->rows.p;
    ab.key = b->key.p; ab.val = b->val.p; ab.key2 = b->key2.p; ab.val2 = b->val2.p;
    ab.tmp = b->tmp.p; ab.tmp_bytes = b->tmp_bytes; ab.seg = b->seg.p;
    ab.G = b->d_grad; ab.C = b->comp.p;
    for (int64_t c = 0; c < b->n_chunk; ++c) {
        const int64_t begin = c * b->chunk;
        const int64_t n_max = std::min(b->chunk, b->hit_capacity - begin);
        if ((e = oa_launch::adjoint_chunk(P, ab, begin, n_max, c == 0 ? 1 : 0, grid, nullptr, s)) != hipSuccess)
            return e;
    }
    return hipSuccess;
}

static int batch_create(mi_sim* sim, int32_t n_event, const int32_t* event_ids, const int64_t* gs_offsets,
                        const void* d_gensteps, void* d_photon, void* d_seq, void* d_hit, int64_t hit_capacity,
                        const mi_adjoint* adj, void* d_grad, int64_t* photon_offsets, mi_batch** out)
{
    if (!out) return fail(MI_ERR_ARG, "null out");
    *out = nullptr;
    if (!sim || n_event <= 0 || !event_ids || !gs_offsets || !photon_offsets || !d_gensteps)
        return fail(MI_ERR_ARG, "bad batch arguments");
    if (d_hit && !d_photon) return fail(MI_ERR_ARG, "hits are gathered from the photon array: pass d_photon");
    for (int32_t e = 0; e < n_event; ++e)
        if (gs_offsets[e + 1] < gs_offsets[e] || gs_offsets[0] != 0) return fail(MI_ERR_ARG, "gs_offsets not ascending from 0");
    const int64_t n_gs = gs_offsets[n_event];
    if (n_gs <= 0) return fail(MI_NOTHING_TO_DO, "zero gensteps");
    mi_batch* b = new mi_batch();
    b->sim = sim;
    b->n_event = n_event;
    b->n_gs = n_gs;
    b->hit_capacity = hit_capacity;
    auto bail = [&](int rc) { mi_batch_destroy(b); return rc; };
#define BCHECK(expr) do { hipError_t e_ = (expr); if (e_ != hipSuccess) return bail(fail(MI_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_))); } while (0)
    hipStream_t s = sim->stream;
    BCHECK(b->prefix.reserve((size_t)n_gs + 1));
    BCHECK(b->saved_prefix.reserve((size_t)n_gs + 1));
    BCHECK(oa_launch::gs_prefix((const float*)d_gensteps, n_gs, b->prefix.p, s));
    std::vector<int64_t> hp((size_t)n_gs + 1);
    BCHECK(hipMemcpyAsync(hp.data(), b->prefix.p, sizeof(int64_t) * (n_gs + 1), hipMemcpyDeviceToHost, s));
    BCHECK(hipMemcpyAsync(b->saved_prefix.p, b->prefix.p, sizeof(int64_t) * (n_gs + 1), hipMemcpyDeviceToDevice, s));
    BCHECK(hipStreamSynchronize(s));
    std::vector<int64_t> evh(2 * ((size_t)n_event + 1));
    for (int e = 0; e <= n_event; ++e) {
        evh[e] = gs_offsets[e];
        evh[n_event + 1 + e] = hp[gs_offsets[e]];
        photon_offsets[e] = hp[gs_offsets[e]];
    }
    b->n_photon = hp[n_gs];
    if (b->n_photon == 0) return bail(fail(MI_NOTHING_TO_DO, "gensteps carry no photons"));
    int64_t max_ev = 0;
    for (int e = 0; e < n_event; ++e) max_ev = std::max(max_ev, photon_offsets[e + 1] - photon_offsets[e]);
    int prc = check_precooked_rows(sim, max_ev);
    if (prc) return bail(prc);
    const uint32_t* xw_table = nullptr;
    prc = xorwow_table(sim, event_ids, n_event, b->xw_jump, &xw_table);
    if (prc) return bail(prc);
    b->n_words = (b->n_photon + 63) / 64;
    const int64_t nb = (b->n_words + 255) / 256;
    BCHECK(upload(b->evbuf, evh.data(), evh.size(), s));
    BCHECK(upload(b->ev_id, event_ids, (size_t)n_event, s));
    BCHECK(b->words.reserve((size_t)b->n_words + 1));
    BCHECK(b->seed.reserve((size_t)b->n_photon));
    BCHECK(b->gs_event.reserve((size_t)n_gs));
    BCHECK(b->next.reserve(1));
    BCHECK(b->flag.reserve(1));
    BCHECK(b->word_prefix.reserve((size_t)b->n_words + 1));
    BCHECK(b->block_total.reserve((size_t)nb + 1));
    BCHECK(b->ho.reserve((size_t)n_event + 1));
    BCHECK(hipHostMalloc((void**)&b->h_out, sizeof(int64_t) * ((size_t)n_event + 2), hipHostMallocDefault));
    for (auto& e : b->ev) BCHECK(hipEventCreate(&e));
    BCHECK(hipStreamCreate(&b->cap));
    SimParams P = base_params(sim);
    P.xw_jump = xw_table;
    P.gensteps = (const float*)d_gensteps;
    P.gs_prefix = b->prefix.p;
    P.n_gs = n_gs;
    P.n_event = n_event;
    P.ev_gs_begin = b->evbuf.p;
    P.ev_ph_begin = b->evbuf.p + (n_event + 1);
    P.ev_id = b->ev_id.p;
    P.n_photon = b->n_photon;
    P.photon_offset = 0;
    P.photon = (float4*)d_photon;
    P.seq = (uint64_t*)d_seq;
    P.hitwords = b->words.p;
    P.next = b->next.p;
    P.seed_map = b->seed.p;
    P.gs_event = b->gs_event.p;
    if (adj) {
        if (!d_hit || hit_capacity <= 0) return bail(fail(MI_ERR_ARG, "the batch adjoint needs the hit buffer"));
        if (!d_grad) return bail(fail(MI_ERR_ARG, "null gradient buffer"));
        int rc = check_adjoint(sim, adj);
        if (rc) return bail(rc);
        b->has_adj = true;
        b->n_mat = adj->n_mat; b->n_bin = adj->n_bin; b->n_det = adj->n_det;
        b->n_param = adj->n_block * 2 * adj->n_bin;
        b->d_grad = (float*)d_grad;
        const int32_t K = sim->cfg.adjoint_tape;
        const bool taping = K > 0 && sim->cfg.kernel == MI_KERNEL_WAVEFRONT && oa_launch::tape_reduce_rows(b->n_param) > 0;
        /* a tape of at least max_bounce records holds every photon (one segment per step):
           no score rows, and one chunk of hits unless the caller limits it */
        const bool tape_only = taping && K >= sim->cfg.max_bounce;
        const int64_t rows_max = adj->max_row_floats > 0 ? adj->max_row_floats : OA_ADJ_ROWS_MAX;
        b->chunk = (tape_only && adj->max_row_floats <= 0)
                       ? std::max<int64_t>(1, std::min<int64_t>(hit_capacity, (int64_t)1 << 30))
                       : std::max<int64_t>(1, std::min<int64_t>(hit_capacity, rows_max / b->n_param));
        b->n_chunk = (hit_capacity + b->chunk - 1) / b->chunk;
        b->tmp_bytes = std::max<size_t>(1, oa_launch::adjoint_sort_bytes(b->chunk, (uint32_t)b->n_det));
        BCHECK(upload(b->mat_block, adj->mat_block, (size_t)adj->n_mat + 1, s));
        BCHECK(upload(b->sample_bin, adj->sample_bin, (size_t)sim->n_wl, s));
        if (b->rows.reserve((size_t)((tape_only ? 1 : b->chunk) * b->n_param)) != hipSuccess)
            return bail(fail(MI_ERR_CAPACITY, "batch adjoint: cannot allocate " +
                                                  std::to_string(b->chunk * b->n_param * 4) + " bytes of score rows"));
        BCHECK(b->comp.reserve((size_t)b->n_det * b->n_param));
        BCHECK(b->key.reserve((size_t)b->chunk));
        BCHECK(b->val.reserve((size_t)b->chunk));
        BCHECK(b->key2.reserve((size_t)b->chunk));
        BCHECK(b->val2.reserve((size_t)b->chunk));
        BCHECK(b->tmp.reserve(b->tmp_bytes));
        BCHECK(b->seg.reserve((size_t)b->n_det + 1));
        BCHECK(hipEventCreate(&b->ev_mid));
        if (taping) {
            if (b->tape.reserve((size_t)b->n_photon * 3 * (size_t)K) != hipSuccess ||
                b->tape_n.reserve((size_t)b->n_photon) != hipSuccess || b->tape_flag.reserve(1) != hipSuccess)
                return bail(fail(MI_ERR_CAPACITY, "batch adjoint: cannot allocate the segment tape (" +
                                                      std::to_string(b->n_photon * 12 * (int64_t)K) + " bytes)"));
            P.tape = b->tape.p; P.tape_n = b->tape_n.p; P.tape_stride = b->n_photon; P.tape_k = K;
            P.tape_overflow = b->tape_flag.p;
        }
    }
    /* one graph: forward child graph, an event record node (splits the
       device time), adjoint child graph */
    hipGraph_t gf = nullptr, ga = nullptr;
    BCHECK(hipStreamBeginCapture(b->cap, hipStreamCaptureModeThreadLocal));
    hipError_t we = batch_work(b, P, d_gensteps, d_photon, d_hit, b->cap);
    hipError_t ce = hipStreamEndCapture(b->cap, &gf);
    if (we != hipSuccess || ce != hipSuccess) {
        if (gf) (void)hipGraphDestroy(gf);
        BCHECK(we);
        BCHECK(ce);
    }
    if (!b->has_adj) {
        b->graph = gf;
    } else {
        BCHECK(hipStreamBeginCapture(b->cap, hipStreamCaptureModeThreadLocal));
        we = batch_adjoint_work(b, P, d_hit, b->cap);
        ce = hipStreamEndCapture(b->cap, &ga);
        if (we != hipSuccess || ce != hipSuccess) {
            (void)hipGraphDestroy(gf);
            if (ga) (void)hipGraphDestroy(ga);
            BCHECK(we);
            BCHECK(ce);
        }
        hipGraph_t parent = nullptr;
        hipGraphNode_t nf = nullptr, nm = nullptr, na = nullptr;
        hipError_t e1 = hipGraphCreate(&parent, 0);
        if (e1 == hipSuccess) e1 = hipGraphAddChildGraphNode(&nf, parent, nullptr, 0, gf);
        if (e1 == hipSuccess) e1 = hipGraphAddEventRecordNode(&nm, parent, &nf, 1, b->ev_mid);
        if (e1 == hipSuccess) e1 = hipGraphAddChildGraphNode(&na, parent, &nm, 1, ga);
        (void)hipGraphDestroy(gf);
        (void)hipGraphDestroy(ga);
        b->graph = parent;
        BCHECK(e1);
    }
    BCHECK(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
#undef BCHECK
    b->generation = sim->generation;
    *out = b;
    return MI_OK;
}

int mi_batch_create(mi_sim* sim, int32_t n_event, const int32_t* event_ids, const int64_t* gs_offsets,
                    const void* d_gensteps, void* d_photon, void* d_seq, void* d_hit, int64_t hit_capacity,
                    int64_t* photon_offsets, mi_batch** out)
{
    return batch_create(sim, n_event, event_ids, gs_offsets, d_gensteps, d_photon, d_seq, d_hit, hit_capacity,
                        nullptr, nullptr, photon_offsets, out);
}

int mi_batch_create_adjoint(mi_sim* sim, int32_t n_event, const int32_t* event_ids, const int64_t* gs_offsets,
                            const void* d_gensteps, void* d_photon, void* d_seq, void* d_hit, int64_t hit_capacity,
                            const mi_adjoint* adj, void* d_grad, int64_t* photon_offsets, mi_batch** out)
{
    if (!adj) return fail(MI_ERR_ARG, "null adjoint description");
    return batch_create(sim, n_event, event_ids, gs_offsets, d_gensteps, d_photon, d_seq, d_hit, hit_capacity, adj,
                        d_grad, photon_offsets, out);
}

int mi_batch_launch(mi_batch* b, void* stream)
{
    if (!b) return fail(MI_ERR_ARG, "null batch");
    if (b->generation != b->sim->generation)
        return fail(MI_ERR_ARG, "the handle's TLAS, precooked uniforms or physics/random-stream settings changed "
                                "since capture: recreate the batch");
    hipStream_t s = stream ? (hipStream_t)stream : b->sim->stream;
    if (b->launched) HIPCHECK(hipEventSynchronize(b->ev[1]));   /* one replay in flight per batch */
    HIPCHECK(hipEventRecord(b->ev[0], s));
    HIPCHECK(hipGraphLaunch(b->exec, s));
    HIPCHECK(hipEventRecord(b->ev[1], s));
    b->launched = true;
    return MI_OK;
}

int mi_batch_wait(mi_batch* b, int64_t* hit_offsets, double* device_ms)
{
    if (!b) return fail(MI_ERR_ARG, "null batch");
    if (!b->launched) return fail(MI_ERR_ARG, "batch not launched");
    HIPCHECK(hipEventSynchronize(b->ev[1]));
    float ms = 0.f;
    HIPCHECK(hipEventElapsedTime(&ms, b->ev[0], b->ev[1]));
    b->last_ms = ms;
    if (device_ms) *device_ms = ms;
    if (b->has_adj) {
        float f = 0.f, a = 0.f;
        HIPCHECK(hipEventElapsedTime(&f, b->ev[0], b->ev_mid));
        HIPCHECK(hipEventElapsedTime(&a, b->ev_mid, b->ev[1]));
        b->last_fwd_ms = f; b->last_adj_ms = a;
    } else {
        b->last_fwd_ms = ms; b->last_adj_ms = 0.0;
    }
    if (hit_offsets)
        for (int32_t e = 0; e <= b->n_event; ++e) hit_offsets[e] = b->h_out[e];
    if (b->h_out[b->n_event + 1] != 0)
        return fail(MI_ERR_ARG, "the gensteps' photon counts differ from those captured: results invalid");
    if (b->hit_capacity > 0 && b->h_out[b->n_event] > b->hit_capacity)
        return fail(MI_ERR_CAPACITY, "hit_capacity too small: " + std::to_string(b->h_out[b->n_event]) + " hits");
    return MI_OK;
}

int64_t mi_batch_num_photon(const mi_batch* b) { return b ? b->n_photon : -1; }

int mi_batch_last_timing(const mi_batch* b, double* ms2)
{
    if (!b || !ms2) return fail(MI_ERR_ARG, "null argument");
    ms2[0] = b->last_fwd_ms;
    ms2[1] = b->last_adj_ms;
    return MI_OK;
}

/* ----------------------------------------------------------- autotuner */

int mi_sim_autotune(mi_sim* sim, int32_t n_event, const int32_t* event_ids, const int64_t* gs_offsets,
                    const void* d_gensteps, void* d_photon, void* d_seq, void* d_hit, int64_t hit_capacity,
                    int32_t reps, mi_tune_result* results, int32_t max_results, int32_t* n_results)
{
    if (!sim || !n_results) return fail(MI_ERR_ARG, "null argument");
    *n_results = 0;
    if (reps < 1) reps = 1;
    static const int blocks[] = { 64, 128, 256 };
    static const int batches[] = { 1, 16, 32 };
    static const int leaves[] = { 1, 2, 4 };
    static const int regroups[] = { 0, 25 };
    const mi_config base = sim->cfg;
    mi_config best = base;
    double best_ms = 1e300;
    std::vector<int64_t> po((size_t)n_event + 1), ho((size_t)n_event + 1);
    int rc = MI_OK;
    for (int leaf : leaves) {
        for (int blk : blocks) {
            for (int rb : batches) {
                for (int rg : regroups) {
                    mi_config c = base;
                    c.bvh_leaf_size = leaf; c.block_size = blk; c.refill_batch = rb; c.trace_regroup_pct = rg;
                    rc = mi_sim_set_config(sim, &c);
                    if (rc) break;
                    rc = mi_sim_simulate_batch_device(sim, n_event, event_ids, gs_offsets, d_gensteps, po.data(),
                                                      d_photon, d_seq, d_hit, hit_capacity, ho.data(), nullptr);
                    if (rc) break;
                    double tot = 0.0;
                    for (int r = 0; r < reps && !rc; ++r) {
                        rc = mi_sim_simulate_batch_device(sim, n_event, event_ids, gs_offsets, d_gensteps, po.data(),
                                                          d_photon, d_seq, d_hit, hit_capacity, ho.data(), nullptr);
                        tot += sim->last_ms[3];
                    }
                    if (rc) break;
                    const double ms = tot / reps;
                    if (results && *n_results < max_results) {
                        mi_tune_result& t = results[*n_results];
                        t.block_size = blk; t.refill_period = c.refill_period; t.bvh_leaf_size = leaf;
                        t.kernel = c.kernel; t.trace_regroup_pct = rg; t.refill_batch = rb; t.ms = ms;
                    }
                    ++*n_results;
                    if (ms < best_ms) { best_ms = ms; best = c; }
                }
                if (rc) break;
            }
            if (rc) break;
        }
        if (rc) break;
    }
    if (rc) {
        const std::string err = g_err;
        (void)mi_sim_set_config(sim, &base);
        g_err = err;
        return rc;
    }
    return mi_sim_set_config(sim, &best);
}

int mi_sim_last_timing(const mi_sim* sim, double* ms4)
{
    if (!sim || !ms4) return fail(MI_ERR_ARG, "null");
    for (int i = 0; i < 4; ++i) ms4[i] = sim->last_ms[i];
    return MI_OK;
}

} // extern "C"
