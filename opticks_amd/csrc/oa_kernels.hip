/*
 * oa_kernels.hip : gfx950 kernels of the propagation path.
 *
 *   k_gs_prefix       exclusive prefix of genstep numphoton (QEvt.cu:181-226 seeding
 *                     reduce)
 *   k_seed_fill       the expand: genstep index of every photon, one block per
 *                     genstep writing its photon range (QEvt.cu:181-226 seed
 *                     buffer, sysrap/iexpand.h); k_gs_event the event of every
 *                     genstep for batched launches
 *   k_sim_plain       one lane per photon, the CSGOptiX7.cu:432-530 shape
 *   k_sim_wave        persistent waves: lanes whose photon finished take a new
 *                     one from a global counter every refill_period bounces
 *                     (wavefront compaction of live photons)
 *   k_hit_count/k_hit_scan/k_hit_scatter
 *                     stable stream compaction of flagmask&hit_mask photons via
 *                     a per-64-photon bitmask (QEvt::gatherHit_, SU.cu:145-160)
 *   k_trace           nearest-hit query (simtrace-style)
 *   k_inst_boxes/k_refit_level  GPU refit of the TLAS after instance moves
 */
#include <hip/hip_runtime.h>
#include "oa_physics.h"
#include "oa_kernels.h"

using namespace oa;

/* Occupancy request for the propagation kernels: at least OA_WAVES_PER_EU
   waves per SIMD caps their VGPR budget (512/N). 0 leaves it to the compiler. */
#ifndef OA_WAVES_PER_EU
#define OA_WAVES_PER_EU 3
#endif
#if OA_WAVES_PER_EU > 0
#define OA_OCC __attribute__((amdgpu_waves_per_eu(OA_WAVES_PER_EU)))
#else
#define OA_OCC
#endif

/* ------------------------------------------------------------ seeding */

/* seed[i] = genstep of photon i for i < n_photon: blocks stride over gensteps,
   threads over each genstep's photon range [prefix[g], prefix[g+1]) (coalesced
   writes). Photons past the gensteps' total (a caller's n_photon larger than
   the gensteps hold) get the last genstep, as the binary search gave them. */
__global__ void __launch_bounds__(256) k_seed_fill(const int64_t* prefix, int64_t n_gs, int64_t n_photon,
                                                   uint32_t* seed)
{
    for (int64_t g = blockIdx.x; g < n_gs; g += gridDim.x) {
        const int64_t b = prefix[g], e = prefix[g + 1] < n_photon ? prefix[g + 1] : n_photon;
        for (int64_t i = b + threadIdx.x; i < e; i += blockDim.x) seed[i] = (uint32_t)g;
    }
    if (blockIdx.x == 0)
        for (int64_t i = prefix[n_gs] + threadIdx.x; i < n_photon; i += blockDim.x) seed[i] = (uint32_t)(n_gs - 1);
}

/* gs_event[g] = event owning genstep g, from the events' genstep boundaries */
__global__ void __launch_bounds__(256) k_gs_event(const int64_t* ev_gs_begin, int32_t n_event, int32_t* gs_event)
{
    for (int32_t e = blockIdx.x; e < n_event; e += gridDim.x)
        for (int64_t g = ev_gs_begin[e] + threadIdx.x; g < ev_gs_begin[e + 1]; g += blockDim.x) gs_event[g] = e;
}

__global__ void __launch_bounds__(1024) k_gs_prefix(const float* gs, int64_t n_gs, int64_t* prefix)
{
    __shared__ int64_t part[1024];
    const int t = threadIdx.x;
    const int64_t per = (n_gs + 1023) / 1024;
    const int64_t b = t * per, e = (b + per < n_gs) ? b + per : n_gs;
    int64_t s = 0;
    for (int64_t g = b; g < e; ++g) s += (int64_t)__float_as_uint(gs[24 * g + 3]);
    part[t] = s;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
        int64_t v = t >= off ? part[t - off] : 0;
        __syncthreads();
        part[t] += v;
        __syncthreads();
    }
    int64_t run = part[t] - s;    /* exclusive */
    for (int64_t g = b; g < e; ++g) { prefix[g] = run; run += (int64_t)__float_as_uint(gs[24 * g + 3]); }
    if (t == 1023) prefix[n_gs] = part[1023];
}

/* --------------------------------------------------------- per photon */

struct Lane {
    Photon p;
    Rng rng;
    uint64_t sh0, sh1, sb0, sb1;
    int bounce;
    int64_t g;
    float lcost, lx, ly;          /* last trace's lposcost and local x, y (lite instantiations) */
};

OA_D void seq_add(Lane& L, unsigned slot, uint32_t flag, uint32_t boundary)
{
    /* sseq::add_nibble (sysrap/sseq.h:174-183) */
    const unsigned iseq = slot / 16u;
    const unsigned shift = 4u * (slot - iseq * 16u);
    const uint64_t f = (uint64_t)((flag ? (uint32_t)__builtin_ffs((int)flag) : 0u) & 0xfu) << shift;
    const uint64_t bn = (uint64_t)(boundary & 0xfu) << shift;
    /* value selects, not a branch per field: a pointer phi would pin Lane to scratch */
    L.sh0 |= iseq == 0u ? f : 0ull;
    L.sb0 |= iseq == 0u ? bn : 0ull;
    L.sh1 |= iseq == 1u ? f : 0ull;
    L.sb1 |= iseq == 1u ? bn : 0ull;
}

OA_D int64_t upper_index(const int64_t* a, int64_t lo, int64_t hi, int64_t v)
{
    /* largest k in [lo, hi) with a[k] <= v */
    while (hi - lo > 1) {
        int64_t mid = (lo + hi) >> 1;
        if (a[mid] <= v) lo = mid; else hi = mid;
    }
    return lo;
}

/* generation for global photon g: genstep lookup, RNG stream, generate_photon */
OA_D void lane_start(const SimParams& P, Lane& L, int64_t g)
{
    L.g = g;
    int e = 0;
    int64_t gsi;
    if (P.seed_map) {
        gsi = (int64_t)P.seed_map[g];
        if (P.n_event > 1) e = P.gs_event[gsi];
    } else {
        if (P.n_event > 1) e = (int)upper_index(P.ev_ph_begin, 0, P.n_event, g);
        gsi = upper_index(P.gs_prefix, P.ev_gs_begin[e], P.ev_gs_begin[e + 1], g);
    }
    const int64_t in_event = g - P.ev_ph_begin[e];
    const uint64_t photon_idx = (uint64_t)(P.photon_offset + in_event);
    L.rng.kind = P.rng_kind;
    if (P.rng_kind == 1) {
        L.rng.cursor = 0;
        const bool ok = (int64_t)photon_idx < P.pre_ni;
        L.rng.pre = ok ? P.precooked + (int64_t)photon_idx * P.pre_nk : nullptr;
        L.rng.pre_nk = ok ? P.pre_nk : 0;
    } else {
        oa_rng_init(&L.rng.ph, P.seed, P.rng_offset, P.skipahead, (uint64_t)P.ev_id[e], photon_idx);
    }
    const float4* gq = reinterpret_cast<const float4*>(P.gensteps + 24 * gsi);
    const int gencode = (int)__float_as_uint(gq[0].x);
    /* qsim::generate_photon dispatch (qudarap/qsim.h:2532-2561) */
    if (gencode == 6) gen_torch(L.p, L.rng, gq, photon_idx);
    else if (gencode == 15 || gencode == 18) gen_cerenkov(P, L.p, L.rng, gq);
    else gen_scint(P, L.p, L.rng, gq);
    /* sphoton::set_index (sysrap/sphoton.h:230-237) */
    L.p.index = (uint32_t)(photon_idx & 0xffffffffu);
    L.p.identity = ((uint32_t)((photon_idx >> 32) & 0xffu) << 24) | (L.p.identity & 0xffffffu);
    L.sh0 = L.sh1 = L.sb0 = L.sb1 = 0ull;
    L.bounce = 0;
    L.lcost = 0.f; L.lx = 0.f; L.ly = __uint_as_float(0x7fc00000u);
    seq_add(L, 0u, ph_flag(L.p), ph_boundary(L.p));
}

/* one iteration of the CSGOptiX7.cu:477-509 loop; returns true when the photon is done */
template <int STRIDE, class Tape, bool LITE = false, bool EXT = false>
OA_D bool lane_step(const SimParams& P, Lane& L, uint32_t* stk, Tape& tape)
{
    if (!(L.bounce < P.max_bounce && L.p.time < P.max_time)) return true;
    const float tmin = (L.p.obf & P.eps0_mask) ? P.tmin0 : P.tmin;
    Prd prd;
    trace<STRIDE, EXT>(P, stk, L.p.pos, L.p.mom, tmin, P.tmax, prd);
    if constexpr (LITE) { L.lcost = prd.lposcost; L.lx = prd.lx; L.ly = prd.ly; }
    if (prd.boundary == 0xffffu) return true;
    prd.normal = normalize(prd.normal);
    OA_COUNT(OA_ST_STEP);
    OA_WAVE_COUNT(OA_ST_W_PHYS, OA_ST_L_PHYS);
    const int command = propagate(P, L.rng, L.p, prd, tape);
    L.bounce++;
    seq_add(L, (unsigned)L.bounce, ph_flag(L.p), ph_boundary(L.p));
    if (command == CMD_BREAK) return true;
    return !(L.bounce < P.max_bounce && L.p.time < P.max_time);
}

/* sphotonlite::init + set_lpos (sysrap/sphotonlite.h:192-198,234-244) from the
   last trace (CSGOptiX7.cu:520-526); lposcost and lposfphi quantized to 16 bits
   with the float->u16 conversion clamped to [0, 0xffff] */
OA_D uint4 make_lite(const Lane& L)
{
    const float fphi = (L.ly != L.ly) ? 0.f : oa_fphi(L.ly, L.lx);
    const float qc = fminf(fmaxf(L.lcost * 65535.f + 0.5f, 0.f), 65535.f);
    const float qf = fminf(fmaxf(fphi * 65535.f + 0.5f, 0.f), 65535.f);
    return make_uint4((1u << 16) | (L.p.identity & 0xffffu), __float_as_uint(L.p.time),
                      ((uint32_t)qc << 16) | (uint32_t)qf, L.p.flagmask);
}

template <bool LITE = false>
OA_D void lane_finish(const SimParams& P, const Lane& L)
{
    const int64_t g = L.g;
    if constexpr (LITE) {
        if (P.photonlite) P.photonlite[g] = make_lite(L);
    }
    if (P.photon) {
        float4* o = P.photon + 4 * g;
        o[0] = make_float4(L.p.pos.x, L.p.pos.y, L.p.pos.z, L.p.time);
        o[1] = make_float4(L.p.mom.x, L.p.mom.y, L.p.mom.z, __uint_as_float(L.p.hitcount_iindex));
        o[2] = make_float4(L.p.pol.x, L.p.pol.y, L.p.pol.z, L.p.wavelength);
        o[3] = make_float4(__uint_as_float(L.p.obf), __uint_as_float(L.p.identity),
                           __uint_as_float(L.p.index), __uint_as_float(L.p.flagmask));
    }
    if (P.seq) {
        ulonglong2* s = reinterpret_cast<ulonglong2*>(P.seq + 4 * g);
        s[0] = make_ulonglong2(L.sh0, L.sh1);
        s[1] = make_ulonglong2(L.sb0, L.sb1);
    }
}

/* ----------------------------------------------------- plain kernel */

/* F: 1 = sphotonlite output, 2 = extended leaves */
template <int BLOCK, int F>
__global__ void __launch_bounds__(BLOCK) OA_OCC k_sim_plain(SimParams P)
{
    __shared__ uint32_t s_stack[OA_STACK_DEPTH * BLOCK];
    uint32_t* stk = s_stack + threadIdx.x;
    const int64_t g = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
    bool hit = false;
    if (g < P.n_photon) {
        Lane L;
        NoTape tape;
        lane_start(P, L, g);
        while (!lane_step<BLOCK, NoTape, (F & 1) != 0, (F & 2) != 0>(P, L, stk, tape)) {}
        lane_finish<(F & 1) != 0>(P, L);
        hit = (L.p.flagmask & P.hit_mask) != 0u;
    }
    const unsigned long long m = __ballot(hit);
    if ((threadIdx.x & 63) == 0 && g < P.n_photon) P.hitwords[g >> 6] = m;
}

/* ------------------------------------------------ persistent kernel */

template <int BLOCK, int F>
__global__ void __launch_bounds__(BLOCK) OA_OCC k_sim_wave(SimParams P)
{
    __shared__ uint32_t s_stack[OA_STACK_DEPTH * BLOCK];
    uint32_t* stk = s_stack + threadIdx.x;
    const int lane = threadIdx.x & 63;
    Lane L;
    NoTape tape;
    bool active = false;
    bool exhausted = false;
    const int period = P.refill_period > 0 ? P.refill_period : 1;
    /* lanes refill together once `batch` of them are idle: photons started in the
       same iteration from one genstep stay in step (the same boundary at the same
       bounce), so the CSG tests of a detector shell run with more lanes at once */
    const int batch = P.refill_batch > 1 ? P.refill_batch : 1;
    for (int iter = 0;; ++iter) {
        OA_CYC_MARK(c_iter);
        const unsigned long long idle = __ballot(!active);
        const unsigned long long all = __ballot(true);
        if (!exhausted && idle != 0ull && (idle == all || ((iter % period) == 0 && __popcll(idle) >= batch))) {
            const int n_idle = __popcll(idle);
            const int leader = __ffsll((long long)idle) - 1;
            unsigned long long base = 0ull;
            if (lane == leader) base = atomicAdd(P.next, (unsigned long long)n_idle);
            base = __shfl(base, leader);
            if (!active) {
                const unsigned long long below = idle & ((1ull << lane) - 1ull);
                const int64_t g = (int64_t)(base + (unsigned long long)__popcll(below));
                if (g < P.n_photon) {
                    lane_start(P, L, g);
                    active = true;
                }
            }
            if ((int64_t)(base + (unsigned long long)n_idle) >= P.n_photon) exhausted = true;
        }
        const unsigned long long act = __ballot(active);
        if (act == 0ull) break;
#ifdef OA_STATS
        if (lane == 0) { OA_COUNT(OA_ST_ITER); OA_COUNT_N(OA_ST_ACTIVE, __popcll(act)); }
#endif
        OA_CYC_ADD(OA_ST_C_REFILL, c_iter);
        if (active) {
            if (lane_step<BLOCK, NoTape, (F & 1) != 0, (F & 2) != 0>(P, L, stk, tape)) {
                lane_finish<(F & 1) != 0>(P, L);
                if ((L.p.flagmask & P.hit_mask) != 0u)
                    atomicOr(P.hitwords + (L.g >> 6), 1ull << (L.g & 63));
                active = false;
            }
        }
        OA_CYC_ADD(OA_ST_C_ITER, c_iter);
    }
}

/* ------------------------------------------------------------ simtrace */

/* CSGOptiX7.cu:563-604 simtrace: one ray per photon of OpticksGenstep_FRAME
   gensteps, generated by qsim::generate_photon_simtrace_frame
   (qudarap/qsim.h:2460-2522) on RNG stream (event 0, photon index), recorded as
   sevent::add_simtrace (sysrap/sevent.h:670-696); a miss writes q0 = (0.6,0.6,0.6,1),
   the miss program's midgrey background (CSGOptiX7.cu:683-700, CSGOptiX/SBT.cc:185-191) */
template <int BLOCK, bool EXT>
__global__ void __launch_bounds__(BLOCK) OA_OCC k_simtrace(SimParams P, float4* out)
{
    __shared__ uint32_t s_stack[OA_STACK_DEPTH * BLOCK];
    const int64_t idx = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
    if (idx >= P.n_photon) return;
    const int64_t gsi = upper_index(P.gs_prefix, 0, P.n_gs, idx);
    const float4* gq = reinterpret_cast<const float4*>(P.gensteps + 24 * gsi);
    Rng rng;
    rng.kind = 0;
    oa_rng_init(&rng.ph, P.seed, P.rng_offset, P.skipahead, 0ull, (uint64_t)idx);
    const uint32_t gridaxes = __float_as_uint(gq[0].y);
    const float u0 = uni(rng);
    float sinPhi, cosPhi;
    oa_sincosf(2.f * OA_PI_F * u0, &sinPhi, &cosPhi);
    const float u1 = uni(rng);
    const float cosTheta = 2.f * u1 - 1.f;
    const float sinTheta = sqrtf(1.f - cosTheta * cosTheta);
    v3 dir = gridaxes == 1u ? mk(0.f, cosPhi, sinPhi)
           : gridaxes == 2u ? mk(cosPhi, 0.f, sinPhi)
           : gridaxes == 3u ? mk(cosPhi, sinPhi, 0.f)
           : mk(sinTheta * cosPhi, sinTheta * sinPhi, cosTheta);
    const float4 q1 = gq[1];
    const v3 pos = q_right(gq + 2, mk(q1.x, q1.y, q1.z), 1.f);
    dir = q_right(gq + 2, dir, 0.f);
    Prd prd;
    trace<BLOCK, EXT>(P, s_stack + threadIdx.x, pos, dir, P.tmin, P.tmax, prd);
    const bool miss = prd.boundary == 0xffffu;
    const float t = miss ? 1.f : prd.t;
    const uint32_t gpb = miss ? 0xffffffffu : (((prd.gprim & 0xffffu) << 16) | (prd.boundary & 0xffffu));
    const uint32_t ii = miss ? 0xffffffffu : prd.iindex_identity;
    float4* o = out + 4 * idx;
    o[0] = miss ? make_float4(0.6f, 0.6f, 0.6f, t) : make_float4(prd.normal.x, prd.normal.y, prd.normal.z, t);
    o[1] = make_float4(pos.x + t * dir.x, pos.y + t * dir.y, pos.z + t * dir.z, P.tmin);
    o[2] = make_float4(pos.x, pos.y, pos.z, __uint_as_float(gpb));
    o[3] = make_float4(dir.x, dir.y, dir.z, __uint_as_float(ii));
}

/* ---------------------------------------------------- hit compaction */

/* per word popcount -> in-block exclusive prefix; block totals */
__global__ void __launch_bounds__(256) k_hit_count(const unsigned long long* words, int64_t n_words,
                                                   int32_t* word_prefix, int64_t* block_total)
{
    __shared__ int32_t s[256];
    const int t = threadIdx.x;
    const int64_t w = (int64_t)blockIdx.x * 256 + t;
    const int32_t c = w < n_words ? __popcll(words[w]) : 0;
    s[t] = c;
    __syncthreads();
    for (int off = 1; off < 256; off <<= 1) {
        int32_t v = t >= off ? s[t - off] : 0;
        __syncthreads();
        s[t] += v;
        __syncthreads();
    }
    if (w < n_words) word_prefix[w] = s[t] - c;
    if (t == 255) block_total[blockIdx.x] = s[255];
}

/* exclusive scan of block totals in place; total into block_total[n_blocks] */
__global__ void __launch_bounds__(1024) k_hit_scan(int64_t* block_total, int64_t n_blocks)
{
    __shared__ int64_t part[1024];
    const int t = threadIdx.x;
    const int64_t per = (n_blocks + 1023) / 1024;
    const int64_t b = t * per, e = (b + per < n_blocks) ? b + per : n_blocks;
    int64_t s = 0;
    for (int64_t i = b; i < e; ++i) s += block_total[i];
    part[t] = s;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
        int64_t v = t >= off ? part[t - off] : 0;
        __syncthreads();
        part[t] += v;
        __syncthreads();
    }
    int64_t run = part[t] - s;
    for (int64_t i = b; i < e; ++i) { int64_t v = block_total[i]; block_total[i] = run; run += v; }
    if (t == 1023) block_total[n_blocks] = part[1023];
}

/* one wave per bitmask word; lane i copies photon 64*w+i when its bit is set */
__global__ void __launch_bounds__(256) k_hit_scatter(const unsigned long long* words, int64_t n_words,
                                                     const int32_t* word_prefix, const int64_t* block_total,
                                                     const float4* photon, float4* hit, int64_t hit_cap)
{
    const int lane = threadIdx.x & 63;
    const int64_t w = ((int64_t)blockIdx.x * 256 + threadIdx.x) >> 6;
    if (w >= n_words) return;
    const unsigned long long m = words[w];
    if (!((m >> lane) & 1ull)) return;
    const int64_t base = block_total[w >> 8] + word_prefix[w];
    const int64_t r = base + __popcll(m & ((1ull << lane) - 1ull));
    if (r >= hit_cap) return;
    const float4* src = photon + 4 * (w * 64 + lane);
    float4* dst = hit + 4 * r;
    dst[0] = src[0]; dst[1] = src[1]; dst[2] = src[2]; dst[3] = src[3];
}

/* hits among photons [0, ph) for each event boundary ph = ev_ph_begin[e] */
__global__ void k_hit_event_offsets(const unsigned long long* words, const int32_t* word_prefix,
                                    const int64_t* block_total, const int64_t* ev_ph_begin, int32_t n_event,
                                    int64_t n_words, int64_t* hit_offsets)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e > n_event) return;
    const int64_t ph = ev_ph_begin[e];
    const int64_t w = ph >> 6;
    int64_t r;
    if (w >= n_words) r = block_total[(n_words + 255) / 256];
    else r = block_total[w >> 8] + word_prefix[w] + __popcll(words[w] & ((1ull << (ph & 63)) - 1ull));
    hit_offsets[e] = r;
}

/* --------------------------------------------------------- trace query */

template <int BLOCK, bool EXT>
__global__ void __launch_bounds__(BLOCK) k_trace(SimParams P, const float4* rays, int64_t n, float4* out)
{
    __shared__ uint32_t s_stack[OA_STACK_DEPTH * BLOCK];
    const int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
    if (i >= n) return;
    const float4 a = rays[2 * i], b = rays[2 * i + 1];
    Prd prd;
    trace<BLOCK, EXT>(P, s_stack + threadIdx.x, mk(a.x, a.y, a.z), mk(b.x, b.y, b.z), a.w, b.w, prd);
    const bool miss = prd.boundary == 0xffffu;
    const v3 nn = miss ? prd.normal : normalize(prd.normal);
    out[2 * i] = make_float4(nn.x, nn.y, nn.z, prd.t);
    const uint32_t inst = miss ? 0xffffffffu : (prd.iindex_identity >> 16);
    out[2 * i + 1] = make_float4(__uint_as_float(prd.iindex_identity), __uint_as_float(prd.boundary),
                                 prd.lposcost, __uint_as_float(inst));
}

/* --------------------------------------------------------------- refit */

/* union of a node's non-empty child boxes */
__device__ static void node_box(const BvhNode& n, float* lo, float* hi)
{
    for (int k = 0; k < 3; ++k) { lo[k] = 1e30f; hi[k] = -1e30f; }
    for (int c = 0; c < OA_BVH_WIDTH; ++c) {
        if (n.child[c] == OA_BVH_EMPTY) continue;
        lo[0] = fminf(lo[0], n.lox[c]); lo[1] = fminf(lo[1], n.loy[c]); lo[2] = fminf(lo[2], n.loz[c]);
        hi[0] = fmaxf(hi[0], n.hix[c]); hi[1] = fmaxf(hi[1], n.hiy[c]); hi[2] = fmaxf(hi[2], n.hiz[c]);
    }
}

/* world box of each instance: transform the 8 corners of its solid's BLAS root box */
__global__ void k_inst_boxes(const float* o2w, const BvhNode* blas, const uint32_t* inst_blas_root,
                             int64_t n_inst, float* inst_box, float pad_rel, float pad_abs)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_inst) return;
    float lo[3], hi[3];
    node_box(blas[inst_blas_root[i]], lo, hi);
    const float* m = o2w + 16 * i;
    float wl[3] = { 1e30f, 1e30f, 1e30f }, wh[3] = { -1e30f, -1e30f, -1e30f };
    for (int c = 0; c < 8; ++c) {
        const float v0 = (c & 1) ? hi[0] : lo[0], v1 = (c & 2) ? hi[1] : lo[1], v2 = (c & 4) ? hi[2] : lo[2];
        for (int k = 0; k < 3; ++k) {
            const float p = v0 * m[k] + v1 * m[4 + k] + v2 * m[8 + k] + m[12 + k];
            wl[k] = fminf(wl[k], p); wh[k] = fmaxf(wh[k], p);
        }
    }
    for (int k = 0; k < 3; ++k) {
        const float pad = pad_rel * (wh[k] - wl[k]) + pad_abs;
        inst_box[6 * i + k] = wl[k] - pad;
        inst_box[6 * i + 3 + k] = wh[k] + pad;
    }
}

__device__ static void link_box(const BvhNode* nodes, const uint32_t* items, const float* inst_box, uint32_t c,
                                float* lo, float* hi)
{
    for (int k = 0; k < 3; ++k) { lo[k] = 1e30f; hi[k] = -1e30f; }
    if (c == OA_BVH_EMPTY) return;
    if (c & OA_BVH_LEAF_BIT) {
        const uint32_t f = c & 0x00ffffffu, n = (c >> 24) & 0x7fu;
        for (uint32_t j = 0; j < n; ++j) {
            const float* b = inst_box + 6 * items[f + j];
            for (int k = 0; k < 3; ++k) { lo[k] = fminf(lo[k], b[k]); hi[k] = fmaxf(hi[k], b[3 + k]); }
        }
    } else {
        node_box(nodes[c], lo, hi);
    }
}

/* refit all TLAS nodes of one depth (deepest level first) */
__global__ void k_refit_level(BvhNode* nodes, int64_t n_nodes, uint32_t depth, const uint32_t* items,
                              const float* inst_box)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_nodes) return;
    BvhNode& n = nodes[i];
    if (n.depth != depth) return;
    for (int c = 0; c < OA_BVH_WIDTH; ++c) {
        if (n.child[c] == OA_BVH_EMPTY) continue;
        float lo[3], hi[3];
        link_box(nodes, items, inst_box, n.child[c], lo, hi);
        n.lox[c] = lo[0]; n.loy[c] = lo[1]; n.loz[c] = lo[2];
        n.hix[c] = hi[0]; n.hiy[c] = hi[1]; n.hiz[c] = hi[2];
    }
}

/* ------------------------------------------------------ host launchers */

namespace oa_launch {

/* histograms (2 x 32) and slow-trace records (up to OA_SLOW_MAX x 3 float4)
   of an OA_STATS build; returns the number of slow traces seen, or -1 */
extern "C" long long oa_stats_dump(unsigned long long* hist64, float* slow, int reset)
{
#ifdef OA_STATS
    unsigned long long n = 0;
    if (hipMemcpyFromSymbol(hist64, HIP_SYMBOL(oa_hist), sizeof(unsigned long long) * 64) != hipSuccess) return -1;
    if (hipMemcpyFromSymbol(&n, HIP_SYMBOL(oa_slow_n), sizeof(n)) != hipSuccess) return -1;
    if (slow && hipMemcpyFromSymbol(slow, HIP_SYMBOL(oa_slow), sizeof(float4) * OA_SLOW_MAX * 3) != hipSuccess) return -1;
    if (reset) {
        static const unsigned long long zh[64] = {};
        const unsigned long long z = 0;
        if (hipMemcpyToSymbol(HIP_SYMBOL(oa_hist), zh, sizeof(zh)) != hipSuccess) return -1;
        if (hipMemcpyToSymbol(HIP_SYMBOL(oa_slow_n), &z, sizeof(z)) != hipSuccess) return -1;
    }
    return (long long)n;
#else
    (void)hist64; (void)slow; (void)reset;
    return -1;
#endif
}

/* counters of an OA_STATS build (zeros otherwise) */
extern "C" int oa_stats_read(unsigned long long* out, int reset)
{
#ifdef OA_STATS
    if (hipMemcpyFromSymbol(out, HIP_SYMBOL(oa_stats), sizeof(unsigned long long) * OA_ST_N) != hipSuccess) return -1;
    if (reset) {
        unsigned long long z[OA_ST_N] = {};
        if (hipMemcpyToSymbol(HIP_SYMBOL(oa_stats), z, sizeof(z)) != hipSuccess) return -1;
    }
#else
    for (int i = 0; i < OA_ST_N; ++i) out[i] = 0ull;
    (void)reset;
#endif
    return 0;
}

hipError_t seed_map(const int64_t* d_prefix, int64_t n_gs, int64_t n_photon, uint32_t* d_seed,
                    const int64_t* d_ev_gs_begin, int32_t n_event, int32_t* d_gs_event, hipStream_t s)
{
    const unsigned gb = (unsigned)std::min<int64_t>(std::max<int64_t>(n_gs, 1), 65535);
    hipLaunchKernelGGL(k_seed_fill, dim3(gb), dim3(256), 0, s, d_prefix, n_gs, n_photon, d_seed);
    if (n_event > 1) {
        const unsigned eb = (unsigned)std::min<int32_t>(n_event, 65535);
        hipLaunchKernelGGL(k_gs_event, dim3(eb), dim3(256), 0, s, d_ev_gs_begin, n_event, d_gs_event);
    }
    return hipGetLastError();
}

hipError_t gs_prefix(const float* d_gs, int64_t n_gs, int64_t* d_prefix, hipStream_t s)
{
    hipLaunchKernelGGL(k_gs_prefix, dim3(1), dim3(1024), 0, s, d_gs, n_gs, d_prefix);
    return hipGetLastError();
}

template <int B>
static hipError_t sim_t(const SimParams& P, int kernel, int grid, hipStream_t s)
{
    if (kernel == 0) {
        const int64_t nb = (P.n_photon + B - 1) / B;
        switch ((P.photonlite ? 1 : 0) | (P.ext_leaves ? 2 : 0)) {
            case 0: hipLaunchKernelGGL((k_sim_plain<B, 0>), dim3((unsigned)nb), dim3(B), 0, s, P); break;
            case 1: hipLaunchKernelGGL((k_sim_plain<B, 1>), dim3((unsigned)nb), dim3(B), 0, s, P); break;
            case 2: hipLaunchKernelGGL((k_sim_plain<B, 2>), dim3((unsigned)nb), dim3(B), 0, s, P); break;
            default: hipLaunchKernelGGL((k_sim_plain<B, 3>), dim3((unsigned)nb), dim3(B), 0, s, P); break;
        }
    } else {
        switch ((P.photonlite ? 1 : 0) | (P.ext_leaves ? 2 : 0)) {
            case 0: hipLaunchKernelGGL((k_sim_wave<B, 0>), dim3((unsigned)grid), dim3(B), 0, s, P); break;
            case 1: hipLaunchKernelGGL((k_sim_wave<B, 1>), dim3((unsigned)grid), dim3(B), 0, s, P); break;
            case 2: hipLaunchKernelGGL((k_sim_wave<B, 2>), dim3((unsigned)grid), dim3(B), 0, s, P); break;
            default: hipLaunchKernelGGL((k_sim_wave<B, 3>), dim3((unsigned)grid), dim3(B), 0, s, P); break;
        }
    }
    return hipGetLastError();
}

template <int B>
static int occ_t()
{
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_sim_wave<B, 0>, B, 0) != hipSuccess) return 0;
    return nb;
}

int wave_blocks_per_cu(int block)
{
    switch (block) {
        case 64: return occ_t<64>();
        case 128: return occ_t<128>();
        case 256: return occ_t<256>();
        default: return 0;
    }
}

hipError_t simulate(const SimParams& P, int kernel, int block, int grid, hipStream_t s)
{
    switch (block) {
        case 64: return sim_t<64>(P, kernel, grid, s);
        case 128: return sim_t<128>(P, kernel, grid, s);
        case 256: return sim_t<256>(P, kernel, grid, s);
        default: return hipErrorInvalidValue;
    }
}

hipError_t compact_hits(const unsigned long long* d_words, int64_t n_words, int32_t* d_word_prefix,
                        int64_t* d_block_total, const float4* d_photon, float4* d_hit, int64_t hit_cap,
                        hipStream_t s)
{
    const int64_t nb = (n_words + 255) / 256;
    if (nb > 0) hipLaunchKernelGGL(k_hit_count, dim3((unsigned)nb), dim3(256), 0, s, d_words, n_words, d_word_prefix, d_block_total);
    hipLaunchKernelGGL(k_hit_scan, dim3(1), dim3(1024), 0, s, d_block_total, nb);
    if (d_hit && d_photon && n_words > 0) {
        const int64_t nt = n_words * 64;
        hipLaunchKernelGGL(k_hit_scatter, dim3((unsigned)((nt + 255) / 256)), dim3(256), 0, s,
                           d_words, n_words, d_word_prefix, d_block_total, d_photon, d_hit, hit_cap);
    }
    return hipGetLastError();
}

hipError_t event_hit_offsets(const unsigned long long* d_words, const int32_t* d_word_prefix,
                             const int64_t* d_block_total, const int64_t* d_ev_ph_begin, int32_t n_event,
                             int64_t n_words, int64_t* d_hit_offsets, hipStream_t s)
{
    hipLaunchKernelGGL(k_hit_event_offsets, dim3((n_event + 1 + 255) / 256), dim3(256), 0, s,
                       d_words, d_word_prefix, d_block_total, d_ev_ph_begin, n_event, n_words, d_hit_offsets);
    return hipGetLastError();
}

__global__ void k_zero_u64(unsigned long long* p, int64_t n)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0ull;
}

__global__ void k_prefix_check(const int64_t* a, const int64_t* b, int64_t n, unsigned long long* flag)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && a[i] != b[i]) flag[0] = 1ull;
}

hipError_t prefix_check(const int64_t* a, const int64_t* b, int64_t n, unsigned long long* flag, hipStream_t s)
{
    if (n > 0) hipLaunchKernelGGL(k_prefix_check, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, a, b, n, flag);
    return hipGetLastError();
}

hipError_t zero_u64(unsigned long long* p, int64_t n, hipStream_t s)
{
    if (n > 0) hipLaunchKernelGGL(k_zero_u64, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, p, n);
    return hipGetLastError();
}

#ifndef OA_SIMTRACE_BLOCK
#define OA_SIMTRACE_BLOCK 64   /* 64: 3.55 ms, 128: 3.71, 256: 3.87 on a 10.1M-ray grid */
#endif

hipError_t simtrace(const SimParams& P, float4* out, hipStream_t s)
{
    constexpr int B = OA_SIMTRACE_BLOCK;
    const unsigned nb = (unsigned)((P.n_photon + B - 1) / B);
    if (P.ext_leaves) hipLaunchKernelGGL((k_simtrace<B, true>), dim3(nb), dim3(B), 0, s, P, out);
    else hipLaunchKernelGGL((k_simtrace<B, false>), dim3(nb), dim3(B), 0, s, P, out);
    return hipGetLastError();
}

hipError_t trace(const SimParams& P, const float4* rays, int64_t n, float4* out, hipStream_t s)
{
    if (P.ext_leaves) hipLaunchKernelGGL((k_trace<128, true>), dim3((unsigned)((n + 127) / 128)), dim3(128), 0, s, P, rays, n, out);
    else hipLaunchKernelGGL((k_trace<128, false>), dim3((unsigned)((n + 127) / 128)), dim3(128), 0, s, P, rays, n, out);
    return hipGetLastError();
}

hipError_t refit(const float* d_o2w, const BvhNode* d_blas, const uint32_t* d_inst_blas_root, int64_t n_inst,
                 float* d_inst_box, BvhNode* d_tlas, int64_t n_tlas, uint32_t max_depth,
                 const uint32_t* d_tlas_items, hipStream_t s)
{
    hipLaunchKernelGGL(k_inst_boxes, dim3((unsigned)((n_inst + 255) / 256)), dim3(256), 0, s,
                       d_o2w, d_blas, d_inst_blas_root, n_inst, d_inst_box, 1e-3f, 1e-2f);
    for (int d = (int)max_depth; d >= 0; --d)
        hipLaunchKernelGGL(k_refit_level, dim3((unsigned)((n_tlas + 255) / 256)), dim3(256), 0, s,
                           d_tlas, n_tlas, (uint32_t)d, d_tlas_items, d_inst_box);
    return hipGetLastError();
}

} // namespace oa_launch

/* ============================================================ adjoint
 *
 * Gradient of expected hits per detector with respect to binned absorption
 * and Rayleigh lengths (no counterpart in the reference). Score-function
 * (likelihood-ratio) estimator: for a detected photon
 *   d/dL log p(path) = sum over segments in the material of
 *       ell/L^2   (survival over the segment length ell)
 *     - 1/L       (on the segment that ends with that process: absorption for
 *                  L_abs -- reemitted photons -- or scattering for L_Rayleigh)
 * chained through the FP32 lerp of the table samples (weights 1-a, a) into the
 * parameter bins. Paths are not stored: each hit photon is replayed from its
 * Philox stream (same code, same bits), and its score row is written to S.
 * Rows are then reduced per detector in hit order with Kahan summation.
 */

struct AdjParams {
    const int32_t* mat_block;   /* (n_mat+1) parameter block per 1-based material index, -1 = none */
    int32_t n_mat;
    const int32_t* sample_bin;  /* (n_wl) parameter bin of each table sample */
    int32_t n_bin;              /* bins per property */
    int32_t n_param;            /* blocks * 2 * n_bin */
};

struct ScoreTape {
    const int32_t* optical;
    int32_t n_wl;
    float wl_low, wl_step;
    AdjParams A;
    float* row;
    __device__ inline __attribute__((always_inline)) void seg(int m1_line, float wl, float La, float Ls, float ell, int ev)
    {
        const int mat = optical[4 * m1_line];
        if (mat < 0 || mat > A.n_mat) return;
        const int blk = A.mat_block[mat];
        if (blk < 0) return;
        float fx = (wl - wl_low) / wl_step;
        float fl = floorf(fx);
        float a = fx - fl;
        int i0, i1;
        wrap_pair((int)fl, n_wl, i0, i1);
        const int b0 = A.sample_bin[i0], b1 = A.sample_bin[i1];
        const float sa = ell / (La * La) - (ev == SEG_ABSORB ? 1.f / La : 0.f);
        const float ss = ell / (Ls * Ls) - (ev == SEG_SCATTER ? 1.f / Ls : 0.f);
        float* r = row + blk * 2 * A.n_bin;
        r[b0] += sa * (1.f - a);
        r[b1] += sa * a;
        r[A.n_bin + b0] += ss * (1.f - a);
        r[A.n_bin + b1] += ss * a;
    }
};

/*
 * Hits processed by one adjoint chunk: hits [begin, begin + n) with
 * n = n_max, or -- batched events, hit count known only on the device --
 * n = min(n_max, min(hit_off[n_ev], hit_cap) - begin), where hit_off (n_ev+1)
 * are the per-event hit offsets of the compaction.
 */
OA_D int64_t chunk_hits(int64_t begin, int64_t n_max, const int64_t* hit_off, int32_t n_ev, int64_t hit_cap)
{
    if (!hit_off) return n_max;
    int64_t n = hit_off[n_ev];
    n = n < hit_cap ? n : hit_cap;
    n -= begin;
    n = n < n_max ? n : n_max;
    return n > 0 ? n : 0;
}

/* launch-local photon index of hit h, or -1: single event (hit_off null) from
   the photon index minus photon_offset; batched events through the hit's event */
OA_D int64_t hit_photon(const SimParams& P, const float4* hits, int64_t h, const int64_t* hit_off, int32_t n_ev)
{
    const float4 q3 = hits[4 * h + 3];
    const uint32_t ident = __float_as_uint(q3.y), index = __float_as_uint(q3.z);
    const int64_t idx = ((int64_t)(ident >> 24) << 32) | (int64_t)index;
    if (!hit_off) {
        const int64_t g = idx - P.photon_offset;
        return (g >= 0 && g < P.n_photon) ? g : -1;
    }
    const int e = (int)upper_index(hit_off, 0, n_ev, h);
    const int64_t g = P.ev_ph_begin[e] + idx;
    return (g >= P.ev_ph_begin[e] && g < P.ev_ph_begin[e + 1]) ? g : -1;
}

/* persistent replay: lanes take the next hit of the chunk from a global
   counter (zeroed before the launch) as soon as their photon is done, as
   k_sim_wave does; the score row of hit h is S + (h - begin) * n_param */
template <int BLOCK, bool EXT>
__global__ void __launch_bounds__(BLOCK) OA_OCC k_adjoint_wave(SimParams P, AdjParams A, const float4* hits,
                                                               int64_t begin, int64_t n_max, const int64_t* hit_off,
                                                               int32_t n_ev, int64_t hit_cap, float* S)
{
    __shared__ uint32_t s_stack[OA_STACK_DEPTH * BLOCK];
    uint32_t* stk = s_stack + threadIdx.x;
    const int lane = threadIdx.x & 63;
    const int64_t n = chunk_hits(begin, n_max, hit_off, n_ev, hit_cap);
    ScoreTape tape;
    tape.optical = P.optical; tape.n_wl = P.n_wl; tape.wl_low = P.wl_low; tape.wl_step = P.wl_step;
    tape.A = A;
    tape.row = S;
    Lane L;
    bool active = false;
    bool exhausted = n <= 0;
    const int batch = P.refill_batch > 1 ? P.refill_batch : 1;
    for (;;) {
        const unsigned long long idle = __ballot(!active);
        const unsigned long long all = __ballot(true);
        if (!exhausted && idle != 0ull && (idle == all || __popcll(idle) >= batch)) {
            const int n_idle = __popcll(idle);
            const int leader = __ffsll((long long)idle) - 1;
            unsigned long long base = 0ull;
            if (lane == leader) base = atomicAdd(P.next, (unsigned long long)n_idle);
            base = __shfl(base, leader);
            if (!active) {
                const unsigned long long below = idle & ((1ull << lane) - 1ull);
                const int64_t i = (int64_t)(base + (unsigned long long)__popcll(below));
                if (i < n) {
                    const int64_t g = hit_photon(P, hits, begin + i, hit_off, n_ev);
                    if (g >= 0) {
                        lane_start(P, L, g);
                        tape.row = S + i * (int64_t)A.n_param;
                        active = true;
                    }
                }
            }
            if ((int64_t)(base + (unsigned long long)n_idle) >= n) exhausted = true;
        }
        const unsigned long long act = __ballot(active);
        if (act == 0ull) {
            if (exhausted) break;
            continue;
        }
        if (active && lane_step<BLOCK, ScoreTape, false, EXT>(P, L, stk, tape)) active = false;
    }
}

/* zero the score rows of the chunk's hits (grid-stride) */
__global__ void k_zero_rows(float* S, int32_t n_param, int64_t begin, int64_t n_max, const int64_t* hit_off,
                            int32_t n_ev, int64_t hit_cap)
{
    const int64_t n = chunk_hits(begin, n_max, hit_off, n_ev, hit_cap) * (int64_t)n_param;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        S[i] = 0.f;
}

/* detector key (sensor_identifier - 1) of each of the chunk's n_max slots;
   non-sensors and slots past the chunk's hits get n_det */
__global__ void k_hit_det(const float4* hits, int64_t begin, int64_t n_max, const int64_t* hit_off, int32_t n_ev,
                          int64_t hit_cap, uint32_t n_det, uint32_t* key, uint32_t* val)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_max) return;
    const int64_t n = chunk_hits(begin, n_max, hit_off, n_ev, hit_cap);
    uint32_t k = n_det;
    if (i < n) {
        const uint32_t id = __float_as_uint(hits[4 * (begin + i) + 3].y) & 0x00ffffffu;
        k = (id >= 1u && id <= n_det) ? id - 1u : n_det;
    }
    key[i] = k;
    val[i] = (uint32_t)i;
}

/* start of each detector's run in the sorted keys */
__global__ void k_det_segments(const uint32_t* key, int64_t n, uint32_t n_det, int64_t* seg)
{
    const int64_t d = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (d > (int64_t)n_det) return;
    int64_t lo = 0, hi = n;
    while (lo < hi) { int64_t mid = (lo + hi) >> 1; if (key[mid] < (uint32_t)d) lo = mid + 1; else hi = mid; }
    seg[d] = lo;
}

#ifndef OA_REDUCE_ROWS
#define OA_REDUCE_ROWS 4
#endif

/* G[d,k] = Kahan sum over the detector's hits (in photon order) of S[h,k];
   with first == 0 the sum and its compensation C continue from an earlier
   chunk of hits (chunks are consecutive in photon order) */
__global__ void __launch_bounds__(256) k_det_reduce(const float* S, const uint32_t* order, const int64_t* seg,
                                                    int32_t n_param, float* G, float* C, int first)
{
    const int d = blockIdx.x;
    const int64_t b = seg[d], e = seg[d + 1];
    for (int k = threadIdx.x; k < n_param; k += blockDim.x) {
        const int64_t o = (int64_t)d * n_param + k;
        float sum = first ? 0.f : G[o], c = first ? 0.f : C[o];
        int64_t i = b;
        /* OA_REDUCE_ROWS rows in flight per step (4, 8 and 16 measure alike), summed in the
           same order as one at a time */
        for (; i + OA_REDUCE_ROWS <= e; i += OA_REDUCE_ROWS) {
            float s[OA_REDUCE_ROWS];
#pragma unroll
            for (int u = 0; u < OA_REDUCE_ROWS; ++u) s[u] = S[(int64_t)order[i + u] * n_param + k];
#pragma unroll
            for (int u = 0; u < OA_REDUCE_ROWS; ++u) {
                const float y = s[u] - c;
                const float t = sum + y;
                c = (t - sum) - y;
                sum = t;
            }
        }
        for (; i < e; ++i) {
            const float y = S[(int64_t)order[i] * n_param + k] - c;
            const float t = sum + y;
            c = (t - sum) - y;
            sum = t;
        }
        G[o] = sum;
        C[o] = c;
    }
}

#include <hipcub/hipcub.hpp>

namespace oa_launch {

/* threads per block of the adjoint replay (64, 128 and 256 measure within 1%) */
#ifndef OA_ADJ_BLOCK
#define OA_ADJ_BLOCK 128
#endif

int adjoint_blocks_per_cu()
{
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_adjoint_wave<OA_ADJ_BLOCK, false>, OA_ADJ_BLOCK, 0) != hipSuccess) return 0;
    return nb;
}

hipError_t adjoint_replay(const SimParams& P, const int32_t* d_mat_block, int32_t n_mat, const int32_t* d_sample_bin,
                          int32_t n_bin, int32_t n_param, const float4* d_hit, int64_t begin, int64_t n_max,
                          const int64_t* d_hit_off, int32_t n_ev, int64_t hit_cap, float* d_S, int grid, hipStream_t s)
{
    AdjParams A;
    A.mat_block = d_mat_block; A.n_mat = n_mat; A.sample_bin = d_sample_bin; A.n_bin = n_bin; A.n_param = n_param;
    if (n_max <= 0) return hipGetLastError();
    hipLaunchKernelGGL(k_zero_rows, dim3((unsigned)std::max(1, grid * 4)), dim3(256), 0, s, d_S, n_param, begin, n_max,
                       d_hit_off, n_ev, hit_cap);
    hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(64), 0, s, P.next, (int64_t)1);
    if (P.ext_leaves)
        hipLaunchKernelGGL((k_adjoint_wave<OA_ADJ_BLOCK, true>), dim3((unsigned)std::max(1, grid)), dim3(OA_ADJ_BLOCK), 0, s, P, A, d_hit,
                           begin, n_max, d_hit_off, n_ev, hit_cap, d_S);
    else
        hipLaunchKernelGGL((k_adjoint_wave<OA_ADJ_BLOCK, false>), dim3((unsigned)std::max(1, grid)), dim3(OA_ADJ_BLOCK), 0, s, P, A, d_hit,
                           begin, n_max, d_hit_off, n_ev, hit_cap, d_S);
    return hipGetLastError();
}

static int det_key_bits(uint32_t n_det)
{
    int end_bit = 1;
    while (end_bit < 32 && (1u << end_bit) <= n_det) ++end_bit;
    return end_bit;
}

hipError_t adjoint_reduce(const float4* d_hit, int64_t begin, int64_t n_max, const int64_t* d_hit_off, int32_t n_ev,
                          int64_t hit_cap, uint32_t n_det, int32_t n_param, const float* d_S, uint32_t* d_key,
                          uint32_t* d_val, uint32_t* d_key2, uint32_t* d_val2, void* d_tmp, size_t tmp_bytes,
                          int64_t* d_seg, float* d_G, float* d_C, int first, hipStream_t s)
{
    if (n_max > 0) {
        hipLaunchKernelGGL(k_hit_det, dim3((unsigned)((n_max + 255) / 256)), dim3(256), 0, s, d_hit, begin, n_max,
                           d_hit_off, n_ev, hit_cap, n_det, d_key, d_val);
        hipError_t e = hipcub::DeviceRadixSort::SortPairs(d_tmp, tmp_bytes, d_key, d_key2, d_val, d_val2,
                                                          (int)n_max, 0, det_key_bits(n_det), s);
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL(k_det_segments, dim3((n_det + 1 + 255) / 256), dim3(256), 0, s, d_key2, n_max, n_det, d_seg);
    hipLaunchKernelGGL(k_det_reduce, dim3(n_det), dim3(256), 0, s, d_S, d_val2, d_seg, n_param, d_G, d_C, first);
    return hipGetLastError();
}

size_t adjoint_sort_bytes(int64_t n_hit, uint32_t n_det)
{
    size_t bytes = 0;
    (void)hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, (uint32_t*)nullptr, (uint32_t*)nullptr,
                                             (uint32_t*)nullptr, (uint32_t*)nullptr, (int)n_hit, 0, det_key_bits(n_det));
    return bytes;
}

} // namespace oa_launch
