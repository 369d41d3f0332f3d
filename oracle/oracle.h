/*
 * oracle.h : CPU restatement of the reference's optical-photon hot path.
 *
 * TEST INFRASTRUCTURE ONLY. This library is the checker that the HIP path is
 * compared against (tests/, __graft_entry__.smoke, bench.py cpu_baseline).
 * Nothing in the product (opticks_amd/, include/) links, loads or calls it.
 *
 * It restates, in plain C99 + OpenMP, one photon per iteration:
 *   generation        qudarap/qsim.h:2532-2561, sysrap/storch.h:192-520,
 *                     qudarap/qscint.h:66-200, qudarap/qcerenkov.h
 *   bounce loop       CSGOptiX/CSGOptiX7.cu:432-530
 *   world trace       optixTrace + __intersection__is/__closesthit__ch/__miss__ms
 *                     (CSGOptiX7.cu:113-218,682-967): nearest hit over all
 *                     instances x prims, here by brute force or a plain BVH
 *   CSG intersect     CSG/csg_intersect_tree.h:277-719, csg_intersect_node.h:392-937
 *                     (list nodes), CSG/csg_intersect_leaf.h:174-324 and leaf headers
 *   physics           qudarap/qsim.h:180-201,433-575,604-866,1001-1208,
 *                     1680-1758,1988-2093,2229-2338; qudarap/qbnd.h:108-223
 *   history           sysrap/sseq.h:174-183, sysrap/sctx.h:134-182
 *   outputs           sphotonlite (sysrap/sphotonlite.h:41-60), simtrace records
 *                     (qudarap/qsim.h:2460-2522, sysrap/sevent.h:670-696)
 *   adjoint rows      score-function terms of the binned lengths along each path
 *
 * The bnd/icdf "textures" are read with FP32 linear interpolation and wrap
 * addressing (qudarap/QTex.cc:293-305) instead of the GPU texture unit.
 * Transcendentals and the Philox streams come from oa_portable.h, which is
 * pinned separately against glibc and the Philox known-answer vectors.
 *
 * How each part is pinned (tests/): CSG intersect by bitwise equality with the
 * reference's own CSG headers compiled under oracle/_ref and by committed
 * golden vectors; generation by the geometry of each torch type, the
 * scintillation and Cerenkov distributions (cone angle, spectrum, yield,
 * polarization, time); boundary/surface/bulk physics by closed-form
 * probabilities, Fresnel coefficients and Rayleigh invariants; simtrace by the
 * grid rules and agreement with trace; adjoint rows by closed forms on a water
 * sphere. Parity unpinned: curand's stream values (this checker and the
 * kernels share their own Philox convention) and the GPU texture unit's 8-bit
 * filtering weights, neither of which can be reproduced offline.
 */
#ifndef OA_ORACLE_H
#define OA_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct oracle_geom {
    const float* node;  int64_t n_node;
    const float* prim;  int64_t n_prim;
    const void*  solid; int64_t n_solid;
    const float* itra;  int64_t n_tran;
    const float* inst;  int64_t n_inst;
    const float* plan;  int64_t n_plan;   /* (n_plan,4) planes of convex polyhedra, may be NULL */
} oracle_geom;

typedef struct oracle_phys {
    const float*   bnd; int32_t n_bnd; int32_t n_wl; float wl_low; float wl_step;
    const int32_t* optical;
    const float*   icdf; int32_t n_icdf_layer; int32_t n_icdf; int32_t icdf_hd_factor;
} oracle_phys;

typedef struct oracle_cfg {
    int32_t  max_bounce;
    float    max_time;
    float    tmin;
    float    tmin0;
    uint32_t eps0_mask;
    float    tmax;
    uint64_t seed;
    uint64_t rng_offset;
    uint64_t event_skipahead;
    int32_t  rng_kind;          /* 0 philox, 1 precooked, 2 xorwow */
    const float* precooked;     /* (precooked_ni, precooked_nk)  */
    int64_t  precooked_ni;
    int64_t  precooked_nk;
    int32_t  traverse;          /* 0 brute force, 1 BVH */
    int32_t  nthreads;          /* OpenMP threads, 0 = default */
} oracle_cfg;

/* Build acceleration data for a geometry; returns an opaque handle. */
void* oracle_world_create(const oracle_geom* g);
void  oracle_world_destroy(void* w);

/*
 * Simulate photons [p_begin, p_end) of the event given by gensteps (n_gs,6,4).
 * photon_out (p_end-p_begin,4,4) f32 and seq_out (p_end-p_begin,2,2) u64 may be
 * NULL. Returns the total number of photons of the gensteps, or <0 on error.
 * photon_offset is added to the in-event photon index to form the absolute
 * RNG subsequence, as CSGOptiX7.cu:442 does with photon_slot_offset.
 */
int64_t oracle_simulate(void* world, const oracle_phys* ph, const oracle_cfg* cfg,
                        int32_t event_idx, const float* gensteps, int64_t n_gs,
                        int64_t photon_offset, int64_t p_begin, int64_t p_end,
                        float* photon_out, uint64_t* seq_out);

/*
 * Adjoint score rows (no counterpart in the reference): replays the in-event
 * photons photons[0..n) and writes, per photon, d log p(path)/d theta for the
 * binned absorption and Rayleigh lengths (rows (n, n_param) f32, n_param =
 * n_block*2*n_bin, block order [absorption bins, Rayleigh bins]).
 * mat_block (n_mat+1) maps 1-based material index to a block (<0 none);
 * sample_bin (n_wl) maps each table sample to its bin. Returns 0 or <0.
 */
int oracle_adjoint_rows(void* world, const oracle_phys* ph, const oracle_cfg* cfg, int32_t event_idx,
                        const float* gensteps, int64_t n_gs, int64_t photon_offset,
                        const int64_t* photons, int64_t n, const int32_t* mat_block, int32_t n_mat,
                        const int32_t* sample_bin, int32_t n_bin, int32_t n_param, float* rows);

/* As oracle_simulate, also writing (p_end-p_begin,4) u32 sphotonlite records
   (lite_out may be NULL). */
int64_t oracle_simulate_ex(void* world, const oracle_phys* ph, const oracle_cfg* cfg,
                           int32_t event_idx, const float* gensteps, int64_t n_gs,
                           int64_t photon_offset, int64_t p_begin, int64_t p_end,
                           float* photon_out, uint64_t* seq_out, uint32_t* lite_out);
float oracle_atan2f(float y, float x);
void  oracle_atan2f_n(const float* y, const float* x, float* out, int64_t n);

/* Simtrace records (n,4,4) of OpticksGenstep_FRAME gensteps (layout as
   mi_sim_simtrace_device). Returns the photon count or <0. */
int64_t oracle_simtrace(void* world, const oracle_cfg* cfg, const float* gensteps, int64_t n_gs, float* out);

/* Nearest hit for n rays; same record layout as mi_sim_trace_device. */
int oracle_trace(void* world, const float* rays, int64_t n, float* isect, int32_t traverse);

/* Single prim intersect in its own frame (CSG/CSGQuery.cc:231-262 shape). */
int oracle_intersect_prim(const oracle_geom* g, int32_t prim_idx, float t_min,
                          const float* ray_origin, const float* ray_direction,
                          float* isect4);

/* The packed Boolean LUT words, for comparison with the reference constants. */
void oracle_packed_lut(uint32_t* acloser3, uint32_t* bcloser3);

/* CSG tree evaluations and loop decisions taken since the last reset (test statistics;
   the first reset switches counting on for the rest of the process). */
void oracle_tree_counts(int64_t* evals, int64_t* loops, int32_t reset);

/* Portable math entry points for the accuracy tests. */
float oracle_logf(float x);
void  oracle_logf_n(const float* x, float* y, int64_t n);
void  oracle_sincosf_n(const float* x, float* s, float* c, int64_t n);
void  oracle_sincosf(float x, float* s, float* c);
void  oracle_philox_block(const uint32_t* ctr4, uint32_t k0, uint32_t k1, uint32_t* out4);
void  oracle_uniforms(uint64_t seed, uint64_t offset, uint64_t skip, uint64_t event_idx,
                      uint64_t photon, int32_t n, float* out);
/* xorwow: the first n raw words of a photon stream (and its state after the
   skip, x0..x4, d, when state6 is not NULL); the jump matrix M^n */
void  oracle_xorwow_words(uint64_t seed, uint64_t offset, uint64_t skip, uint64_t event_idx,
                          uint64_t photon, int32_t n, uint32_t* out, uint32_t* state6);
void  oracle_xorwow_jump(uint64_t n, uint32_t* J800);
/* from an explicit state (x0..x4, d): skip `skip` words by the jump matrix, then n words */
void  oracle_xorwow_from_state(const uint32_t* state6, uint64_t skip, int32_t n, uint32_t* out,
                               uint32_t* state_out);

/* Single physics steps for closed-form checks (qudarap/QSim.cu test kernels). */
int oracle_step_boundary(const float* mom, const float* pol, const float* normal,
                         float n1, float n2, float u_reflect,
                         float* mom_out, float* pol_out, uint32_t* flag_out,
                         float* transcoeff_out);
void oracle_rayleigh(const float* mom, const float* pol, uint64_t seed, uint64_t photon,
                     int32_t n, float* mom_out, float* pol_out);

#ifdef __cplusplus
}
#endif

#endif
