/*
 * oa_layout.h : records shared by the host runtime and the gfx950 kernels.
 *
 * HBM layout (all 16-byte aligned, read with dwordx4 loads):
 *   BvhNode  128 B  four child boxes in SoA order (lo.x[4], lo.y[4], lo.z[4],
 *                   hi.x[4], hi.y[4], hi.z[4]) + four child links: one node
 *                   is seven dwordx4 loads and tests four children. Link: bit31
 *                   set = leaf with items [link & 0xffffff, +((link>>24)&0x7f));
 *                   0xffffffff = empty slot; otherwise the child node index.
 *   InstRec   64 B  world->object 3x4 (rows of qat4::right_multiply: m[0..2],
 *                   m[4..6], m[8..10] linear, m[12..14] translation), then
 *                   gas (solid) index, identity (sensor_identifier+1), blas root
 *                   (front instances: one leaf link over all the solid's prims).
 *   PrimRec   16 B  node offset of the prim's root CSGNode, boundary, globalPrimIdx,
 *                   prim index. The BLAS leaves index a copy of the records in leaf
 *                   order, so a prim test reads its record without an item lookup.
 *   nodes/itra      the reference CSGNode / qat4 arrays unchanged (64 B each).
 *   bnd             (nbnd*4*2 rows, nwl, float4): row 2*line+k is one property
 *                   group of one line, contiguous in wavelength so a lerp reads
 *                   32 adjacent bytes.
 */
#pragma once
#include <stdint.h>

#define OA_BVH_LEAF_BIT 0x80000000u
#define OA_BVH_EMPTY    0xffffffffu
#define OA_STACK_DEPTH  48
#define OA_BVH_WIDTH    4
#define OA_CSG_MAX_HEIGHT 7
#define OA_MAX_FRONT    4       /* instances tested before the TLAS walk (world-sized boxes) */

/* Scratch of the no-allocation simulate entry point: one statically sized
   device array per GPU (oa_kernels.hip), laid out as 256-byte aligned pieces.
   Limits per call: OA_NOALLOC_MAX_GS gensteps, OA_NOALLOC_MAX_PHOTON photons. */
#define OA_NOALLOC_MAX_GS      (1ll << 20)
#define OA_NOALLOC_MAX_PHOTON  (1ll << 26)
#define OA_ALIGN256(x)         ((((long long)(x)) + 255ll) & ~255ll)
#define OA_NOALLOC_WORDS       (OA_NOALLOC_MAX_PHOTON / 64 + 1)                 /* hit bitmask words */
#define OA_NOALLOC_BLOCKS      ((OA_NOALLOC_WORDS + 255) / 256 + 1)             /* compaction block totals */
#define OA_NOALLOC_OFF_PREFIX  0ll                                              /* (MAX_GS+1) int64 */
#define OA_NOALLOC_OFF_EV      (OA_NOALLOC_OFF_PREFIX + OA_ALIGN256(8 * (OA_NOALLOC_MAX_GS + 1)))
#define OA_NOALLOC_OFF_EVID    (OA_NOALLOC_OFF_EV + 256ll)                      /* 4 int64 event bounds */
#define OA_NOALLOC_OFF_NEXT    (OA_NOALLOC_OFF_EVID + 256ll)                    /* 1 int32 event id */
#define OA_NOALLOC_OFF_XW      (OA_NOALLOC_OFF_NEXT + 256ll)                    /* 1 u64 refill counter */
#define OA_NOALLOC_OFF_WORDS   (OA_NOALLOC_OFF_XW + OA_ALIGN256(4 * 800))       /* xorwow jump matrix */
#define OA_NOALLOC_OFF_WPREFIX (OA_NOALLOC_OFF_WORDS + OA_ALIGN256(8 * OA_NOALLOC_WORDS))
#define OA_NOALLOC_OFF_BTOTAL  (OA_NOALLOC_OFF_WPREFIX + OA_ALIGN256(4 * OA_NOALLOC_WORDS))
#define OA_NOALLOC_SCRATCH_BYTES (OA_NOALLOC_OFF_BTOTAL + OA_ALIGN256(8 * OA_NOALLOC_BLOCKS))

/* Optional work counters (build with -DOA_STATS): traces, TLAS node visits,
   instance tests, BLAS node visits, prim tests, tree prim tests, leaf
   evaluations inside trees, propagation steps, wave loop iterations, active
   lanes summed over wave loop iterations, and for the trace entry, the TLAS
   walk, the leaf (instance/BLAS) walk, prim tests and physics the number of
   wave executions and the lanes active in them. Off in normal builds. */
enum { OA_ST_TRACE, OA_ST_TLAS, OA_ST_INST, OA_ST_BLAS, OA_ST_PRIM, OA_ST_TREE, OA_ST_LEAF, OA_ST_STEP,
       OA_ST_ITER, OA_ST_ACTIVE,
       /* wave-level pairs: loop iterations executed by the wave, and lanes active in them */
       OA_ST_W_TRACE, OA_ST_L_TRACE, OA_ST_W_P1, OA_ST_L_P1, OA_ST_W_P2, OA_ST_L_P2,
       OA_ST_W_PRIM, OA_ST_L_PRIM, OA_ST_W_PHYS, OA_ST_L_PHYS,
       /* wave-level shader-clock cycles spent in each part (first active lane adds) */
       OA_ST_C_ITER, OA_ST_C_REFILL, OA_ST_C_TRACE, OA_ST_C_P1, OA_ST_C_P2, OA_ST_C_PRIM,
       /* CSG tree evaluations: wave executions, lanes in them, cycles */
       OA_ST_W_TREE, OA_ST_L_TREE, OA_ST_C_TREE,
       /* cycles in the leaf intersections of tree evaluations */
       OA_ST_C_TLEAF,
       /* node visits whose children all miss or lie beyond the nearest hit */
       OA_ST_TLAS_DEAD, OA_ST_BLAS_DEAD,
       /* tree evaluations run by the general postorder algorithm */
       OA_ST_TREE_GENERAL, OA_ST_N };

struct BvhNode {
    float lox[4], loy[4], loz[4];
    float hix[4], hiy[4], hiz[4];
    uint32_t child[4];
    uint32_t depth, pad[3];
};

struct InstRec {
    float m[16];        /* w2o: m[0..2], m[4..6], m[8..10] linear rows, m[12..14] translation;
                           m[3] = gas (as bits), m[7] = identity (bits), m[11] = blas root (bits) */
};

struct PrimRec {
    uint32_t node_offset;
    uint32_t boundary;
    uint32_t global_prim;
    uint32_t index;               /* the prim's index (in the BLAS-ordered copy) */
};

struct SimParams {
    /* geometry */
    const float4*  node;
    const float4*  itra;
    const float4*  plan;          /* convex polyhedron planes (n.xyz, d) */
    int32_t        ext_leaves;    /* geometry holds cone/polyhedron/hyperboloid/phicut/halfspace leaves */
    const InstRec* inst;
    const PrimRec* prim;
    const BvhNode* tlas;
    const uint32_t* tlas_items;   /* instance indices */
    const BvhNode* blas;
    const PrimRec* blas_prims;    /* prim records in BLAS leaf order (PrimRec.index: the prim) */
    uint32_t front_leaf;          /* leaf link to the front instances' items (tested first), or empty */
    int32_t tlas_empty;           /* the TLAS holds no instance: every instance is front */
    const float4* free_ball;      /* (cx, cy, cz, r^2): a ball no TLAS instance box meets (r^2 <= 0: none) */
    /* physics */
    const float4*  bnd;
    const int32_t* optical;       /* (nbnd*4, 4) */
    const float*   icdf;
    int32_t n_wl;
    float   wl_low, wl_step;
    int32_t n_icdf, n_icdf_layer, icdf_hd;
    /* config */
    int32_t  max_bounce;
    float    max_time, tmin, tmin0, tmax;
    uint32_t eps0_mask;
    uint32_t hit_mask;
    uint64_t seed, rng_offset, skipahead;
    int32_t  rng_kind;            /* 0 Philox, 1 precooked, 2 xorwow */
    const float* precooked;       /* (pre_ni, pre_nk) uniforms, a row per photon index */
    int64_t  pre_ni, pre_nk;
    const uint32_t* xw_jump;      /* xorwow: (n_event, 800) jump matrices of each event's skip */
    /* event(s): photons are numbered globally over all events of the launch */
    const float*   gensteps;      /* (n_gs,6,4) all events concatenated */
    const int64_t* gs_prefix;     /* (n_gs+1) exclusive prefix of numphoton */
    const uint32_t* seed_map;     /* (n_photon) genstep of each photon, or null: binary search of gs_prefix */
    const int32_t* gs_event;      /* (n_gs) event of each genstep (with seed_map and n_event > 1) */
    int64_t        n_gs;
    int32_t        n_event;
    const int64_t* ev_gs_begin;   /* (n_event+1) genstep index boundaries */
    const int64_t* ev_ph_begin;   /* (n_event+1) global photon boundaries */
    const int32_t* ev_id;         /* (n_event) event index for skipahead */
    int64_t        n_photon;      /* total photons of the launch */
    int64_t        photon_offset; /* added to the in-event index (single-event slicing) */
    /* outputs */
    float4*        photon;        /* (n_photon,4) float4 or null */
    uint64_t*      seq;           /* (n_photon,2,2) or null */
    uint4*         photonlite;    /* (n_photon) sphotonlite or null (lite kernel instantiation) */
    unsigned long long* hitwords; /* (n_photon+63)/64 bitmask, zeroed before launch */
    /* persistent kernel */
    unsigned long long* next;     /* global photon counter */
    int32_t refill_period;
    int32_t refill_batch;         /* refill once this many lanes are idle (or all are) */
    int32_t regroup_pct;          /* trace: leaf phase once fewer than this % of the walking lanes still walk */
    /* segment records for a following adjoint (null: none), segment-major so that
       lanes at the same step write neighbouring records: record j < tape_k of
       launch-local photon g is the 3 words (bnd line | event << 30, wavelength,
       length) at tape + 3 * (j * tape_stride + g); tape_n[g] is the number of
       segments the photon had (above tape_k: the records stop at tape_k) */
    uint32_t* tape;
    uint8_t*  tape_n;
    int64_t   tape_stride;
    int32_t   tape_k;
    uint32_t* tape_overflow;      /* set when a hit's photon had more segments than the tape holds */
};
