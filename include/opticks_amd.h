/*
 * opticks_amd.h : C-ABI of the MI355X optical-photon propagation path.
 *
 * This is the drop-in boundary for the reference's simulate call:
 * gensteps in, photons / per-photon history (seq) / hits out.
 * It replaces, one for one:
 *
 *   reference                                          here
 *   ---------------------------------------------------------------------------
 *   SSimulator (sysrap/SSimulator.h) implemented by     mi_sim_* (a handle that
 *     CSGOptiX (CSGOptiX/CSGOptiX.cc:349,755,1218)        owns geometry, physics
 *                                                         tables and device buffers)
 *   CSGOptiX::Create(CSGFoundry*) +                     mi_sim_create
 *     QSim::UploadComponents (qudarap/QSim.cc:119)
 *   NP* QSim::simulate(const NP* gs, int eventID)       mi_sim_simulate (host arrays)
 *     (qudarap/QSim.cc:720-755)
 *   QSim::simulate(eventID, reset) per-slice launches   mi_sim_simulate_device
 *     (qudarap/QSim.cc:436-536, CSGOptiX.cc:1091-1189)    (device-resident arrays)
 *   evt->photonlite output (CSGOptiX7.cu:520-526)       mi_sim_simulate_lite_device
 *   SSimulator::reset(eventID) (G4CXOpticks.cc:491-511)  mi_sim_reset
 *   CSGOptiX simtrace (CSGOptiX7.cu:563-604)            mi_sim_simtrace_device
 *   nearest-hit query (CSGOptiX7.cu:113-218)             mi_sim_trace_device
 *   (no counterpart: batched events in one launch)      mi_sim_simulate_batch_device
 *   QBnd upload of bnd (qudarap/QBnd.cc:130-175)         mi_sim_set_bnd (new values,
 *                                                         same shape, no rebuild)
 *   (no counterpart: differentiation)                   mi_sim_adjoint_device
 *   (no counterpart: replayable captured batches)       mi_batch_* (hipGraph)
 *   (no counterpart: launch tuning)                     mi_sim_autotune
 *   SPM::merge_partial_select (sysrap/SPM.cu:135-240)   mi_sim_merge_hits_device
 *
 * Array layouts are the reference's, so arrays exported by the reference
 * (CSGFoundry / SSim folders) can be passed unchanged:
 *   genstep  (G,6,4)  f32/i32  quad6: storch (sysrap/storch.h:42-74) or
 *                                     sscint (sysrap/sscint.h:34-62)
 *   photon   (N,4,4)  f32/u32  sphoton (sysrap/sphoton.h:171-246)
 *   seq      (N,2,2)  u64      [seqhis[2], seqbnd[2]] (sysrap/sseq.h:54-66,
 *                              sysrap/SEvt.cc:3754-3756)
 *   hit      (M,4,4)  f32/u32  photons with flagmask & hit_mask, photon order
 *                              (qudarap/QEvt.cc:1113-1132, sysrap/SU.cu:145-160)
 *   node     (n,4,4)  CSGNode  (CSG/CSGNode.h:196); prims are a leaf, a complete
 *                     Boolean tree, or a list node (CSG_CONTIGUOUS <= 8 leaves,
 *                     CSG_DISCONTIGUOUS, CSG_OVERLAP) whose leaves sit at
 *                     root + subOffset; list nodes may also sit in a tree.
 *                     Leaves: sphere, zsphere, cylinder, box3, cone,
 *                     convexpolyhedron, hyperboloid, phicut, halfspace; other
 *                     typecodes fail create with MI_ERR_UNSUPPORTED
 *   prim     (n,4,4)  CSGPrim  (CSG/CSGPrim.h:75)
 *   solid    (n) 48 B CSGSolid (CSG/CSGSolid.h:37)
 *   itra     (n,4,4)  qat4 inverse node transforms (sysrap/sqat4.h)
 *   plan     (n,4)    convex polyhedron planes (nx,ny,nz,d), addressed by
 *                     CSGNode planeIdx/planeNum (CSG/CSGNode.h:204-207)
 *   inst     (n,4,4)  qat4 instance transforms, identity in the w column
 *                     (sysrap/sqat4.h:345-390)
 *   bnd      (nbnd,4,2,nwl,4) f32 boundary table (qudarap/QBnd.cc:130-175)
 *   optical  (nbnd,4,4) i32 (sysrap/sstandard.h make_optical)
 *   icdf     (nlayer, nicdf) f32 scintillation inverse CDF (qudarap/QScint.cc:25-52)
 *
 * Errors: no exceptions cross this boundary. Functions return MI_OK (0),
 * MI_NOTHING_TO_DO (1, e.g. zero gensteps, cf. QEvt.cc:436 rc=1) or a negative
 * MI_ERR_* code; mi_last_error() gives a thread-local message.
 * Threading: a handle is not re-entrant; the caller serializes calls on it.
 * Several handles may share one device. All device work of a handle runs on
 * the HIP stream given to the call (NULL: the handle's own stream).
 */
#ifndef OPTICKS_AMD_H
#define OPTICKS_AMD_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MI_OK               0
#define MI_NOTHING_TO_DO    1
#define MI_ERR_ARG         -1
#define MI_ERR_HIP         -2
#define MI_ERR_UNSUPPORTED -3
#define MI_ERR_CAPACITY    -4
#define MI_ERR_GEOMETRY    -5

/* output_mask bits */
#define MI_OUT_PHOTON 0x1u
#define MI_OUT_SEQ    0x2u
#define MI_OUT_HIT    0x4u

/* rng_kind: the per-photon random stream (qudarap/qrng.h:68-148, sysrap/srng.h:30-36) */
#define MI_RNG_PHILOX    0    /* Philox4x32-10, counter keyed on absolute photon index */
#define MI_RNG_PRECOOKED 1    /* uniforms read from a caller array (aligned comparison):
                                 row = absolute photon index, the cursor wrapping modulo
                                 precooked_nk (sysrap/s_seq.h:105-129); a launch whose
                                 photons lack a row fails with MI_ERR_ARG */
#define MI_RNG_XORWOW    2    /* Marsaglia xorwow (the RNG_XORWOW generator of qrng.h:68-106),
                                 own seeding per photon and GF(2) jump-ahead for the
                                 offset + event skipahead (oa_portable.h) */

/* kernel: which propagation kernel runs the bounce loop */
#define MI_KERNEL_PLAIN      0   /* one thread per photon, CSGOptiX7.cu:432-530 shape */
#define MI_KERNEL_WAVEFRONT  1   /* persistent waves that refill finished lanes */

typedef struct mi_geometry {
    const float*   node;  int64_t n_node;
    const float*   prim;  int64_t n_prim;
    const void*    solid; int64_t n_solid;
    const float*   itra;  int64_t n_tran;
    const float*   plan;  int64_t n_plan;
    const float*   inst;  int64_t n_inst;
} mi_geometry;

typedef struct mi_physics {
    const float*   bnd;            /* (n_bnd,4,2,n_wl,4)            */
    int32_t        n_bnd;
    int32_t        n_wl;           /* 761 for the 60..820 nm domain  */
    float          wl_low;         /* 60 nm  (sysrap/sdomain.h:26-31) */
    float          wl_step;        /* 1 nm                            */
    const int32_t* optical;        /* (n_bnd,4,4)                     */
    const float*   icdf;           /* (n_icdf_layer, n_icdf) or NULL  */
    int32_t        n_icdf_layer;   /* 1 or 3                          */
    int32_t        n_icdf;         /* 4096                            */
    int32_t        icdf_hd_factor; /* 0, 10 or 20 (qudarap/qscint.h:160-200) */
} mi_physics;

/*
 * Run configuration: names and defaults follow SEventConfig
 * (sysrap/SEventConfig.cc:41-110); mi_config_default fills them in.
 */
typedef struct mi_config {
    int32_t  max_bounce;       /* OPTICKS_MAX_BOUNCE         31      */
    float    max_time;         /* OPTICKS_MAX_TIME           1e27 ns */
    float    tmin;             /* OPTICKS_PROPAGATE_EPSILON  0.05 mm */
    float    tmin0;            /* OPTICKS_PROPAGATE_EPSILON0 0.05 mm */
    uint32_t eps0_mask;        /* OPTICKS_PROPAGATE_EPSILON0_MASK TO,CK,SI,SC,RE */
    float    tmax;             /* 1e6 mm (CSGOptiX/CSGOptiX.cc:662) */
    uint32_t hit_mask;         /* OPTICKS_HIT_MASK "SD"           */
    uint64_t seed;             /* QRng__SEED_OFFSET seed   (0)    */
    uint64_t rng_offset;       /* QRng__SEED_OFFSET offset (0)    */
    uint64_t event_skipahead;  /* OPTICKS_EVENT_SKIPAHEAD 100000  */
    int32_t  rng_kind;         /* MI_RNG_*                        */
    const float* precooked;    /* (precooked_ni, precooked_nk) uniforms, host; copied to the
                                  device by create/set_config. NULL with the ni/nk the handle
                                  already holds keeps that copy (mi_sim_get_config returns NULL) */
    int64_t  precooked_ni;
    int64_t  precooked_nk;     /* 1 .. 2^31-1 */
    uint32_t output_mask;      /* MI_OUT_* bits                   */
    int64_t  max_slot;         /* photons per launch, 0 = from free VRAM */
    int32_t  kernel;           /* MI_KERNEL_*                     */
    int32_t  block_size;       /* threads per block: 64 (default), 128 or 256 */
    int32_t  refill_period;    /* wavefront kernel: bounces between lane refills */
    int32_t  bvh_leaf_size;    /* instances per TLAS leaf (1..8)  */
    int32_t  waves_per_cu;     /* wavefront kernel: resident waves per CU (0 = auto) */
    int32_t  trace_regroup_pct; /* trace: a wave's node walk hands over to the leaf phase once
                                   fewer than this percent of its walking lanes still walk
                                   (0 = once none do; 0..100; results do not depend on it) */
    int32_t  refill_batch;     /* wavefront kernel: a wave refills once this many of its lanes
                                   are idle, or all are (1..64, default 16; results do not depend on it) */
    int32_t  adjoint_tape;     /* wavefront forwards keep up to this many segment records (12 B)
                                   per photon for a following adjoint, which then scores hits from
                                   them and replays only photons with more segments (0..64, default
                                   0 = none: the adjoint replays every hit; gradients are the same) */
} mi_config;

typedef struct mi_sim mi_sim;

/* Host-side results of mi_sim_simulate, owned by the handle until the next
   simulate/reset/destroy (the reference keeps the SEvt fold until reset). */
typedef struct mi_event_out {
    int64_t         n_photon;
    int64_t         n_hit;
    const float*    photon;     /* (n_photon,4,4) or NULL when not in output_mask */
    const uint64_t* seq;        /* (n_photon,2,2) or NULL */
    const float*    hit;        /* (n_hit,4,4)    or NULL */
    double          device_ms;  /* seeding -> hit compaction, HIP events */
} mi_event_out;

void        mi_config_default(mi_config* cfg);
const char* mi_last_error(void);
const char* mi_version(void);

int  mi_sim_create(const mi_geometry* geo, const mi_physics* phys,
                   const mi_config* cfg, mi_sim** out);
void mi_sim_destroy(mi_sim* sim);
/* set_config validates and builds (TLAS for a new leaf size, precooked upload)
   before it commits: on failure the handle keeps its previous configuration.
   A change of any physics or random-stream setting (max_bounce, max_time, tmin,
   tmin0, eps0_mask, tmax, hit_mask, seed, rng_offset, event_skipahead, rng_kind,
   precooked values) or of the TLAS invalidates captured batches (mi_batch_launch
   then fails; recreate them). */
int  mi_sim_set_config(mi_sim* sim, const mi_config* cfg);
int  mi_sim_get_config(const mi_sim* sim, mi_config* cfg);

/* gensteps (n_gs,6,4) in host memory; results copied to handle-owned host buffers. */
int  mi_sim_simulate(mi_sim* sim, int32_t event_id, const float* gensteps,
                     int64_t n_gs, mi_event_out* out);
int  mi_sim_reset(mi_sim* sim, int32_t event_id);

/*
 * Device-resident form. d_gensteps: (n_gs,6,4) in device memory.
 * n_photon: photons to propagate, at most the gensteps' total (a smaller
 * count propagates that prefix), or -1 for the total; the total is read back
 * (8 bytes) and a larger n_photon fails with MI_ERR_ARG.
 * d_photon (n_photon,4,4), d_seq (n_photon,2,2) u64, d_hit (hit_capacity,4,4):
 * caller-owned device buffers; any may be NULL to skip that output.
 * photon_offset: absolute index of the first photon (multi-launch slicing,
 * CSGOptiX/Params.h photon_slot_offset). *n_hit_out gets the hit count.
 * stream: hipStream_t or NULL for the handle's stream. Synchronous on return.
 */
int  mi_sim_simulate_device(mi_sim* sim, int32_t event_id,
                            const void* d_gensteps, int64_t n_gs,
                            int64_t n_photon, int64_t photon_offset,
                            void* d_photon, void* d_seq,
                            void* d_hit, int64_t hit_capacity, int64_t* n_hit_out,
                            void* stream);

/*
 * No counterpart in the reference: the form of mi_sim_simulate_device for callers
 * that cannot allocate (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc). Same arguments, results and errors as
 * mi_sim_simulate_device, but every device buffer beyond the caller's inputs and
 * outputs (genstep prefix, event bounds, refill counter, xorwow jump matrix,
 * hit bitmask, compaction prefixes) lives in one statically sized scratch array
 * per GPU (about 20 MiB, held until the process ends); the call makes no
 * device or host allocation on its success path. Limits: at most
 * MI_NOALLOC_MAX_GENSTEPS gensteps and MI_NOALLOC_MAX_PHOTONS photons per call
 * (MI_ERR_CAPACITY beyond); no max_slot slicing. All handles and calls on one
 * GPU share that scratch: calls that overlap on one GPU (other threads or
 * streams) race and must be serialized by the caller.
 */
#define MI_NOALLOC_MAX_GENSTEPS (1ll << 20)
#define MI_NOALLOC_MAX_PHOTONS  (1ll << 26)
int  mi_sim_simulate_device_noalloc(mi_sim* sim, int32_t event_id,
                                    const void* d_gensteps, int64_t n_gs,
                                    int64_t n_photon, int64_t photon_offset,
                                    void* d_photon, void* d_seq,
                                    void* d_hit, int64_t hit_capacity, int64_t* n_hit_out,
                                    void* stream);

/*
 * Same, also writing d_photonlite (n_photon) sphotonlite records
 * (sysrap/sphotonlite.h:41-60), as CSGOptiX7.cu:520-526 does for
 * evt->photonlite: hitcount 1 | identity, time, 16-bit lposcost / lposfphi of
 * the last trace's local hit position, flagmask. d_photon may be NULL (the lite
 * records are the output), in which case d_hit must be NULL too; lite hits are
 * selected and merged with mi_sim_merge_hits_device.
 */
int  mi_sim_simulate_lite_device(mi_sim* sim, int32_t event_id,
                                 const void* d_gensteps, int64_t n_gs,
                                 int64_t n_photon, int64_t photon_offset,
                                 void* d_photonlite, void* d_photon, void* d_seq,
                                 void* d_hit, int64_t hit_capacity, int64_t* n_hit_out,
                                 void* stream);

/*
 * Batched form: n_event events, event e owns gensteps
 * [gs_offsets[e], gs_offsets[e+1]) of d_gensteps and photons
 * [photon_offsets[e], photon_offsets[e+1]) of d_photon/d_seq, all propagated
 * by one persistent launch. Each event keeps its own RNG streams
 * (event_ids[e] skipahead, photon index from 0 within the event), so results
 * equal per-event mi_sim_simulate_device calls. gs_offsets/event_ids/
 * photon_offsets are host arrays; photon_offsets is filled by the call.
 * Hits of all events are compacted in event order into d_hit;
 * hit_offsets (host, n_event+1) receives the per-event ranges.
 */
int  mi_sim_simulate_batch_device(mi_sim* sim, int32_t n_event,
                                  const int32_t* event_ids,
                                  const int64_t* gs_offsets,
                                  const void* d_gensteps,
                                  int64_t* photon_offsets,
                                  void* d_photon, void* d_seq,
                                  void* d_hit, int64_t hit_capacity,
                                  int64_t* hit_offsets,
                                  void* stream);

/*
 * Nearest-hit query over the whole geometry for n rays, the trace() +
 * closest-hit contract of CSGOptiX7.cu:113-218,776-874:
 * d_rays  (n,2,4) f32: (pos.xyz, tmin), (dir.xyz, tmax)
 * d_isect (n,2,4):     (world normal.xyz normalized, t), then u32
 *                      (iindex<<16|identity, boundary (0xffff = miss),
 *                       lposcost as f32 bits, instance index)
 */
int  mi_sim_trace_device(mi_sim* sim, const void* d_rays, int64_t n,
                         void* d_isect, void* stream);

/*
 * Simtrace (CSGOptiX7.cu:563-604, the reference's SRG_SIMTRACE raygen): one
 * nearest-hit ray per photon of OpticksGenstep_FRAME gensteps (gencode 17;
 * q0 = (gencode, gridaxes, ., numphoton), q1.xyz = local origin, q2..q5 the
 * local->world qat4), directions from qsim::generate_photon_simtrace_frame
 * (qudarap/qsim.h:2460-2522) with RNG stream (event 0, photon index).
 * d_simtrace receives (n_photon,4,4) f32 sevent::add_simtrace records
 * (sysrap/sevent.h:670-696):
 *   q0 world normal.xyz, t       q1 pos + t dir, tmin
 *   q2 pos, u32 globalPrimIdx<<16 | boundary
 *   q3 dir, u32 iindex<<16 | identity
 * A miss writes q0 = (0.6, 0.6, 0.6, 1), the miss program's midgrey
 * background (CSGOptiX/SBT.cc:185-191), with both u32 words 0xffffffff.
 * n_photon < 0 takes the count from the gensteps.
 */
int  mi_sim_simtrace_device(mi_sim* sim, const void* d_gensteps, int64_t n_gs,
                            int64_t n_photon, void* d_simtrace, void* stream);

/* Host form of the same (QSim::simtrace, qudarap/QSim.cc:876-910): gensteps in host
   memory; *simtrace points at (n_simtrace,4,4) records owned by the handle
   until the next simtrace/reset/destroy; *device_ms (may be NULL) the device time. */
int  mi_sim_simtrace(mi_sim* sim, const float* gensteps, int64_t n_gs,
                     const float** simtrace, int64_t* n_simtrace, double* device_ms);

/* Replace instance transforms (n_inst,4,4) and refit the TLAS on the GPU. */
int  mi_sim_refit_instances(mi_sim* sim, const float* inst, int64_t n_inst);

/* Last launch timings in ms (HIP events): [seed, propagate, compact, total]. */
int  mi_sim_last_timing(const mi_sim* sim, double* ms4);

/* Replace the boundary table values (host array, the shape given at create). */
int  mi_sim_set_bnd(mi_sim* sim, const float* bnd);

/*
 * Binned length parameters for the adjoint (no counterpart in the reference).
 * theta is n_block blocks of [n_bin absorption lengths, n_bin Rayleigh lengths];
 * the bnd table sample i of a material in block b holds theta[b][sample_bin[i]]
 * (the caller writes those values with mi_sim_set_bnd).
 * mat_block (host, n_mat+1): block of each 1-based material index, -1 = none.
 * sample_bin (host, n_wl): bin of each wavelength sample, in [0, n_bin).
 * n_det: detectors, hits carry sensor identifiers 1..n_det (others ignored).
 * max_row_floats: device scratch for score rows per chunk of hits, in floats
 * (0 = 2^30, i.e. 4 GiB); hits are processed in chunks that fit it. With a
 * segment tape of at least max_bounce records (mi_config.adjoint_tape) no rows
 * are used, and the hits are chunked only when this is set.
 */
typedef struct mi_adjoint {
    const int32_t* mat_block;
    int32_t n_mat;
    const int32_t* sample_bin;
    int32_t n_bin;
    int32_t n_block;
    int32_t n_det;
    int64_t max_row_floats;
} mi_adjoint;

/*
 * Path-replay adjoint of one event simulated with mi_sim_simulate_device
 * (same handle, config, bnd values, gensteps, event_id, n_photon and
 * photon_offset). d_hit (n_hit,4,4) are that call's hits. Each hit photon is
 * replayed from its RNG stream and its score
 *   d log p(path)/d L = sum over segments ell/L^2 - [segment ends in the process]/L
 * is spread over the bins of the two table samples each lookup interpolates.
 * d_grad (n_det, n_block*2*n_bin) f32 receives d E[hits of detector d]/d theta,
 * summed per detector in photon order with Kahan compensation.
 * When the handle's last forward was this event made with mi_config.adjoint_tape
 * = K > 0 (wavefront kernel), hits are scored from its segment records instead,
 * inside the per-detector reduction, and only photons with more than K segments
 * are replayed; the gradient is the same bit for bit.
 * Device scratch: min(n_hit*n_param, max_row_floats) floats of score rows (none
 * with a tape of K >= max_bounce records), n_det*n_param floats of Kahan
 * compensation, sort buffers.
 * Synchronous on return; mi_sim_last_adjoint_timing gives [replay, reduce, total] ms
 * (with the tape: [sort and overflow replay, fused scoring and reduction, total]).
 */
int  mi_sim_adjoint_device(mi_sim* sim, int32_t event_id,
                           const void* d_gensteps, int64_t n_gs,
                           int64_t n_photon, int64_t photon_offset,
                           const void* d_hit, int64_t n_hit,
                           const mi_adjoint* adj, void* d_grad, void* stream);
int  mi_sim_last_adjoint_timing(const mi_sim* sim, double* ms3);

/*
 * Hit merging, SPM::merge_partial_select (sysrap/SPM.cu:135-240, sysrap/SPM.hh):
 * records d_in (n_in) of type MI_RECORD_PHOTON (sphoton, 64 B) or
 * MI_RECORD_PHOTONLITE (sphotonlite, 16 B, sysrap/sphotonlite.h:41-60) are
 * selected by (flagmask & select_mask) != 0 (MI_MERGE_ALREADY_SELECTED skips
 * the selection), then records with equal key (identity << 48 | (unsigned)(time /
 * time_window)) are merged with the type's reduce_op (sphoton.h:292-320,
 * sphotonlite.h:68-97): earliest time (first in input order on ties), OR of
 * flagmasks, sum of hitcounts. Output in key order into d_out (caller-owned,
 * out_capacity records); time_window 0 returns the selection in input order
 * unmerged. *n_out gets the record count; MI_ERR_CAPACITY if it exceeds
 * out_capacity. Synchronous on return.
 */
#define MI_RECORD_PHOTON          0
#define MI_RECORD_PHOTONLITE      1
#define MI_MERGE_ALREADY_SELECTED 0xffffffffu
int  mi_sim_merge_hits_device(mi_sim* sim, int32_t record, const void* d_in, int64_t n_in,
                              uint32_t select_mask, float time_window,
                              void* d_out, int64_t out_capacity, int64_t* n_out, void* stream);

/*
 * Captured batch (no counterpart in the reference): the batched form's
 * launch sequence -- genstep prefix, persistent propagation, hit compaction,
 * per-event hit offsets -- recorded once into a hipGraph and replayed by
 * mi_batch_launch. Arguments as mi_sim_simulate_batch_device; the device
 * buffers must stay valid while the batch exists. Between launches the
 * gensteps' contents may change but not their per-genstep photon counts
 * (checked on the device; mi_batch_wait then fails with MI_ERR_ARG).
 * Launch settings (kernel, block, refill, waves) are those of the handle at
 * create; a later change of TLAS leaf size or of a physics or random-stream
 * setting (see mi_sim_set_config) invalidates the batch (launch fails;
 * recreate it). bnd values set with mi_sim_set_bnd
 * are seen by later launches. One launch in flight per batch.
 */
typedef struct mi_batch mi_batch;
int  mi_batch_create(mi_sim* sim, int32_t n_event, const int32_t* event_ids,
                     const int64_t* gs_offsets, const void* d_gensteps,
                     void* d_photon, void* d_seq, void* d_hit, int64_t hit_capacity,
                     int64_t* photon_offsets, mi_batch** out);
/*
 * Same, with the path-replay adjoint of every hit of the batch captured into
 * the same graph after the forward part (hipGraph child graphs joined by an
 * event node): d_grad (adj->n_det, n_param) receives d E[hits per detector]/d theta
 * summed over all events of the batch. hit_capacity also bounds the adjoint's
 * hit slots; score rows use min(hit_capacity, max_row_floats/n_param) hits per
 * chunk, and the chunk loop is fixed at capture, reading the actual hit count on
 * the device. With mi_config.adjoint_tape = K > 0 the batch keeps its own segment
 * tape (12 K bytes per photon) and scores hits from it (K >= max_bounce: no score
 * rows, one chunk unless max_row_floats is set). mi_batch_last_timing gives
 * [forward, adjoint] ms of the last replay.
 */
int  mi_batch_create_adjoint(mi_sim* sim, int32_t n_event, const int32_t* event_ids,
                             const int64_t* gs_offsets, const void* d_gensteps,
                             void* d_photon, void* d_seq, void* d_hit, int64_t hit_capacity,
                             const mi_adjoint* adj, void* d_grad,
                             int64_t* photon_offsets, mi_batch** out);
int  mi_batch_launch(mi_batch* b, void* stream);                 /* asynchronous */
int  mi_batch_wait(mi_batch* b, int64_t* hit_offsets, double* device_ms);
int64_t mi_batch_num_photon(const mi_batch* b);
int  mi_batch_last_timing(const mi_batch* b, double* ms2);
int  mi_batch_destroy(mi_batch* b);

/*
 * Autotuner (no counterpart in the reference): times the batched form on the
 * given events for every block size {64,128,256} x refill batch {1,16,32} x
 * TLAS leaf size {1,2,4} x trace regroup percent {0,25} (refill period as
 * configured), reps timed runs each after one warm run,
 * and leaves the handle configured with the fastest. results (optional,
 * max_results entries) receives each candidate's mean device ms.
 */
typedef struct mi_tune_result {
    int32_t block_size, refill_period, bvh_leaf_size, kernel, trace_regroup_pct, refill_batch;
    double ms;
} mi_tune_result;
int  mi_sim_autotune(mi_sim* sim, int32_t n_event, const int32_t* event_ids,
                     const int64_t* gs_offsets, const void* d_gensteps,
                     void* d_photon, void* d_seq, void* d_hit, int64_t hit_capacity,
                     int32_t reps, mi_tune_result* results, int32_t max_results,
                     int32_t* n_results);

#ifdef __cplusplus
}
#endif

#endif /* OPTICKS_AMD_H */
