/*
 * oa_sim.h : the per-photon propagation kernels, templated on the block size,
 * a feature word F (1 = sphotonlite output, 2 = extended leaves and list
 * nodes) and the random generator R (PhiloxRng, XorwowRng, PrecookedRng), so
 * every instantiation carries only the live state its features need.
 *
 *   k_sim_plain       one lane per photon, the CSGOptiX7.cu:432-530 shape
 *   k_sim_wave        persistent waves: lanes whose photon finished take a new
 *                     one from a global counter (wavefront compaction of live
 *                     photons), in batches so photons of one genstep start together
 *   k_adjoint_wave    path replay of hit photons with score terms (no
 *                     counterpart in the reference)
 *
 * oa_sim.hip instantiates them for one (block, generator) pair per compile.
 */
#pragma once
#include <hip/hip_runtime.h>
#include "oa_physics.h"

namespace oa {

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

/* ------------------------------------------------------- lane state */

/* last trace's lposcost and local x, y: kept only by the lite instantiations */
template <bool LITE> struct LiteState { float lcost, lx, ly; };
template <> struct LiteState<false> {};

template <class R, bool LITE>
struct Lane {
    Photon p;
    R rng;
    uint64_t sh0, sh1, sb0, sb1;
    int bounce;
    int64_t g;
    LiteState<LITE> lite;
};

template <class R, bool LITE>
OA_D void seq_add(Lane<R, LITE>& L, unsigned slot, uint32_t flag, uint32_t boundary)
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

/* ------------------------------------------ per-photon random streams */

/* stream of absolute photon index photon_idx in event slot e of the launch:
   subsequence = photon index, skip = offset + event_id * skipahead words
   (qudarap/qrng.h:84-89,131-137) */
OA_D void rng_start(const SimParams& P, PhiloxRng& r, int e, uint64_t photon_idx)
{
    oa_rng_init(&r.ph, P.seed, P.rng_offset, P.skipahead, (uint64_t)P.ev_id[e], photon_idx);
}

OA_D void rng_start(const SimParams& P, PhiloxRngInline& r, int e, uint64_t photon_idx)
{
    oa_rng_init(&r.ph, P.seed, P.rng_offset, P.skipahead, (uint64_t)P.ev_id[e], photon_idx);
}

OA_D void rng_start(const SimParams& P, XorwowRng& r, int e, uint64_t photon_idx)
{
    oa_xorwow_seed(&r.xw, P.seed, photon_idx);
    const uint64_t n = P.rng_offset + P.skipahead * (uint64_t)P.ev_id[e];
    oa_xorwow_skip(&r.xw, P.xw_jump + (int64_t)e * OA_XORWOW_JUMP_WORDS, n);
}

/* the host checks that every photon of the launch has its row */
OA_D void rng_start(const SimParams& P, PrecookedRng& r, int, uint64_t photon_idx)
{
    r.row = P.precooked + (int64_t)photon_idx * P.pre_nk;
    r.cursor = 0;
    r.nk = (int32_t)P.pre_nk;
}

/* generation for global photon g: genstep lookup, RNG stream, generate_photon */
template <class R, bool LITE>
OA_D void lane_start(const SimParams& P, Lane<R, LITE>& L, int64_t g)
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
    rng_start(P, L.rng, e, photon_idx);
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
    if constexpr (LITE) { L.lite.lcost = 0.f; L.lite.lx = 0.f; L.lite.ly = __uint_as_float(0x7fc00000u); }
    seq_add(L, 0u, ph_flag(L.p), ph_boundary(L.p));
}

/* one iteration of the CSGOptiX7.cu:477-509 loop; returns true when the photon is done */
template <int STRIDE, class Tape, bool LITE, bool EXT, class R>
OA_D bool lane_step(const SimParams& P, Lane<R, LITE>& L, uint32_t* stk, Tape& tape)
{
    if (!(L.bounce < P.max_bounce && L.p.time < P.max_time)) return true;
    const float tmin = (L.p.obf & P.eps0_mask) ? P.tmin0 : P.tmin;
    Prd prd;
    trace<STRIDE, EXT>(P, stk, L.p.pos, L.p.mom, tmin, P.tmax, prd);
    if constexpr (LITE) { L.lite.lcost = prd.lposcost; L.lite.lx = prd.lx; L.lite.ly = prd.ly; }
    if (prd.boundary == 0xffffu) return true;
    prd.normal = normalize(prd.normal);
    OA_COUNT(OA_ST_STEP);
    OA_WAVE_COUNT(OA_ST_W_PHYS, OA_ST_L_PHYS);
    tape.at(P, L.g, L.bounce);
    const int command = propagate(P, L.rng, L.p, prd, tape);
    L.bounce++;
    seq_add(L, (unsigned)L.bounce, ph_flag(L.p), ph_boundary(L.p));
    if (command == CMD_BREAK) return true;
    return !(L.bounce < P.max_bounce && L.p.time < P.max_time);
}

/* sphotonlite::init + set_lpos (sysrap/sphotonlite.h:192-198,234-244) from the
   last trace (CSGOptiX7.cu:520-526); lposcost and lposfphi quantized to 16 bits
   with the float->u16 conversion clamped to [0, 0xffff] */
template <class R>
OA_D uint4 make_lite(const Lane<R, true>& L)
{
    const float fphi = (L.lite.ly != L.lite.ly) ? 0.f : oa_fphi(L.lite.ly, L.lite.lx);
    const float qc = fminf(fmaxf(L.lite.lcost * 65535.f + 0.5f, 0.f), 65535.f);
    const float qf = fminf(fmaxf(fphi * 65535.f + 0.5f, 0.f), 65535.f);
    return make_uint4((1u << 16) | (L.p.identity & 0xffffu), __float_as_uint(L.p.time),
                      ((uint32_t)qc << 16) | (uint32_t)qf, L.p.flagmask);
}

template <class R, bool LITE>
OA_D void lane_finish(const SimParams& P, const Lane<R, LITE>& L)
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

} // namespace oa

/* ----------------------------------------------------- plain kernel */

template <int BLOCK, int F, class R>
__global__ void __launch_bounds__(BLOCK) OA_OCC k_sim_plain(SimParams P)
{
    using namespace oa;
    __shared__ uint32_t s_stack[OA_STACK_DEPTH * BLOCK];
    uint32_t* stk = s_stack + threadIdx.x;
    const int64_t g = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
    bool hit = false;
    if (g < P.n_photon) {
        Lane<R, (F & 1) != 0> L;
        NoTape tape;
        lane_start(P, L, g);
        while (!lane_step<BLOCK, NoTape, (F & 1) != 0, (F & 2) != 0>(P, L, stk, tape)) {}
        lane_finish(P, L);
        hit = (L.p.flagmask & P.hit_mask) != 0u;
    }
    const unsigned long long m = __ballot(hit);
    if ((threadIdx.x & 63) == 0 && g < P.n_photon) P.hitwords[g >> 6] = m;
}

/* ------------------------------------------------ persistent kernel */

template <int BLOCK, int F, class R, class T = oa::NoTape>
__global__ void __launch_bounds__(BLOCK) OA_OCC k_sim_wave(SimParams P)
{
    using namespace oa;
    __shared__ uint32_t s_stack[OA_STACK_DEPTH * BLOCK];
    uint32_t* stk = s_stack + threadIdx.x;
    const int lane = threadIdx.x & 63;
    Lane<R, (F & 1) != 0> L;
    T tape;
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
            if (lane_step<BLOCK, T, (F & 1) != 0, (F & 2) != 0>(P, L, stk, tape)) {
                lane_finish(P, L);
                tape.finish(P, L.g, L.bounce);
                if ((L.p.flagmask & P.hit_mask) != 0u)
                    atomicOr(P.hitwords + (L.g >> 6), 1ull << (L.g & 63));
                active = false;
            }
        }
        OA_CYC_ADD(OA_ST_C_ITER, c_iter);
    }
}

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
 * random stream (same code, same bits), and its score row is written to S.
 * Rows are then reduced per detector in hit order with Kahan summation.
 */

struct AdjParams {
    const int32_t* mat_block;   /* (n_mat+1) parameter block per 1-based material index, -1 = none */
    int32_t n_mat;
    const int32_t* sample_bin;  /* (n_wl) parameter bin of each table sample */
    int32_t n_bin;              /* bins per property */
    int32_t n_param;            /* blocks * 2 * n_bin */
};

/* the four score-row updates of one segment, in this order:
   row[i0] += sa (1 - a), row[i1] += sa a, row[i2] += ss (1 - a), row[i3] += ss a
   (i0, i1 need not differ: two table samples can fall in one bin) */
struct ScoreTerm {
    int32_t i0, i1, i2, i3;
    float sa, ss, a;
    bool on;
};

struct ScoreTape {
    const int32_t* optical;
    int32_t n_wl;
    float wl_low, wl_step;
    AdjParams A;
    float* row;
    __device__ inline __attribute__((always_inline)) void at(const SimParams&, int64_t, int) {}
    /* score terms of a segment in material line m1_line at wavelength wl, length ell,
       ending with event ev, where the table gives absorption length La and Rayleigh
       length Ls: d log p / dL = ell / L^2 - [event of that process] / L, split over
       the parameter bins of the two lerped table samples with weights 1 - a, a */
    __device__ inline __attribute__((always_inline)) ScoreTerm term(int m1_line, float wl, float La, float Ls, float ell,
                                                                    int ev) const
    {
        using namespace oa;
        ScoreTerm t;
        t.on = false;
        t.i0 = t.i1 = t.i2 = t.i3 = 0;
        t.sa = t.ss = t.a = 0.f;
        const int mat = optical[4 * m1_line];
        if (mat < 0 || mat > A.n_mat) return t;
        const int blk = A.mat_block[mat];
        if (blk < 0) return t;
        float fx = (wl - wl_low) / wl_step;
        float fl = floorf(fx);
        float a = fx - fl;
        int i0, i1;
        wrap_pair((int)fl, n_wl, i0, i1);
        const int b0 = A.sample_bin[i0], b1 = A.sample_bin[i1];
        const float sa = ell / (La * La) - (ev == SEG_ABSORB ? 1.f / La : 0.f);
        const float ss = ell / (Ls * Ls) - (ev == SEG_SCATTER ? 1.f / Ls : 0.f);
        const int o = blk * 2 * A.n_bin;
        t.i0 = o + b0;
        t.i1 = o + b1;
        t.i2 = o + A.n_bin + b0;
        t.i3 = o + A.n_bin + b1;
        t.sa = sa; t.ss = ss; t.a = a;
        t.on = true;
        return t;
    }
    static __device__ inline __attribute__((always_inline)) void apply(float* r, const ScoreTerm& t)
    {
        if (!t.on) return;
        r[t.i0] += t.sa * (1.f - t.a);
        r[t.i1] += t.sa * t.a;
        r[t.i2] += t.ss * (1.f - t.a);
        r[t.i3] += t.ss * t.a;
    }
    __device__ inline __attribute__((always_inline)) void seg(const SimParams&, int m1_line, float wl, float La, float Ls,
                                                              float ell, int ev)
    {
        apply(row, term(m1_line, wl, La, Ls, ell, ev));
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
    const int e = (int)oa::upper_index(hit_off, 0, n_ev, h);
    const int64_t g = P.ev_ph_begin[e] + idx;
    return (g >= P.ev_ph_begin[e] && g < P.ev_ph_begin[e + 1]) ? g : -1;
}

/* persistent replay: lanes take the next hit of the chunk from a global
   counter (zeroed before the launch) as soon as their photon is done, as
   k_sim_wave does; the score row of hit h is S + (h - begin) * n_param. With
   the segment tape only hits whose photon had more segments than it holds are
   replayed, and their rows are read by the fused tape reduction. */
template <int BLOCK, bool EXT, class R>
__global__ void __launch_bounds__(BLOCK) OA_OCC k_adjoint_wave(SimParams P, AdjParams A, const float4* hits,
                                                               int64_t begin, int64_t n_max, const int64_t* hit_off,
                                                               int32_t n_ev, int64_t hit_cap, float* S)
{
    using namespace oa;
    __shared__ uint32_t s_stack[OA_STACK_DEPTH * BLOCK];
    uint32_t* stk = s_stack + threadIdx.x;
    const int lane = threadIdx.x & 63;
    /* with the segment tape, only photons it could not hold are replayed: none, when
       the ordering pass (k_hit_det) raised no flag */
    if (P.tape && *P.tape_overflow == 0u) return;
    const int64_t n = chunk_hits(begin, n_max, hit_off, n_ev, hit_cap);
    ScoreTape tape;
    tape.optical = P.optical; tape.n_wl = P.n_wl; tape.wl_low = P.wl_low; tape.wl_step = P.wl_step;
    tape.A = A;
    tape.row = S;
    Lane<R, false> L;
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
                    /* a hit whose segments are all on the tape was scored from it */
                    if (g >= 0 && !(P.tape_n && (int32_t)P.tape_n[g] <= P.tape_k)) {
                        lane_start(P, L, g);
                        tape.row = S + i * (int64_t)A.n_param;
                        /* with the tape the rows are not cleared beforehand: only these hits use them */
                        if (P.tape)
                            for (int32_t k = 0; k < A.n_param; ++k) tape.row[k] = 0.f;
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

/* threads per block of the adjoint replay (64, 128 and 256 measure within 1%) */
#ifndef OA_ADJ_BLOCK
#define OA_ADJ_BLOCK 128
#endif
