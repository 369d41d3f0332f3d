/* oa_kernels.h : host-side launchers of the gfx950 kernels (oa_kernels.hip). */
#pragma once
#include <hip/hip_runtime.h>
#include "oa_layout.h"

namespace oa_launch {
/* genstep of each photon (seed) and, for n_event > 1, event of each genstep */
hipError_t seed_map(const int64_t* d_prefix, int64_t n_gs, int64_t n_photon, uint32_t* d_seed,
                    const int64_t* d_ev_gs_begin, int32_t n_event, int32_t* d_gs_event, hipStream_t s);
hipError_t gs_prefix(const float* d_gs, int64_t n_gs, int64_t* d_prefix, hipStream_t s);
hipError_t simulate(const SimParams& P, int kernel, int block, int grid, hipStream_t s);
int wave_blocks_per_cu(int block, int rng_kind);   /* resident blocks of the persistent kernel per CU */
hipError_t prefix_check(const int64_t* a, const int64_t* b, int64_t n, unsigned long long* flag, hipStream_t s);
hipError_t zero_u64(unsigned long long* p, int64_t n, hipStream_t s);
hipError_t compact_hits(const unsigned long long* d_words, int64_t n_words, int32_t* d_word_prefix,
                        int64_t* d_block_total, const float4* d_photon, float4* d_hit, int64_t hit_cap,
                        hipStream_t s);
hipError_t event_hit_offsets(const unsigned long long* d_words, const int32_t* d_word_prefix,
                             const int64_t* d_block_total, const int64_t* d_ev_ph_begin, int32_t n_event,
                             int64_t n_words, int64_t* d_hit_offsets, hipStream_t s);
hipError_t simtrace(const SimParams& P, float4* out, hipStream_t s);
hipError_t trace(const SimParams& P, const float4* rays, int64_t n, float4* out, hipStream_t s);
hipError_t refit(BvhNode* d_tlas, int64_t n_tlas, uint32_t max_depth, const uint32_t* d_tlas_items,
                 const float* d_inst_box, hipStream_t s);
/* device buffers of the adjoint: parameter map (mat_block, sample_bin), hits with
   optional per-event hit offsets (batched events: the hit count is read on the
   device and clipped to hit_cap), score rows S (n_max rows of n_param), sort
   keys/values and temporary storage, detector runs seg (n_det + 1), and the
   per-detector sums G with their Kahan compensation C (n_det x n_param) */
struct AdjBufs {
    const int32_t* mat_block; int32_t n_mat;
    const int32_t* sample_bin; int32_t n_bin; int32_t n_param;
    const float4* hit; const int64_t* hit_off; int32_t n_ev; int64_t hit_cap;
    uint32_t n_det;
    float* S;
    uint32_t *key, *val, *key2, *val2;
    void* tmp; size_t tmp_bytes;
    int64_t* seg;
    float *G, *C;
};
/* adjoint of hits [begin, begin+n) (n = n_max, or the device count): score rows,
   then per-detector Kahan sums in photon order (first: start them, else continue
   from the previous chunk). Without P.tape every hit is replayed into dense rows;
   with it, hits are scored from their tape records inside the per-detector
   reduction and only photons the tape could not hold are replayed. `mid` (may be
   null) is recorded before the reduction. */
hipError_t adjoint_chunk(const SimParams& P, const AdjBufs& a, int64_t begin, int64_t n_max, int first, int grid,
                         hipEvent_t mid, hipStream_t s);
/* hits per round of the fused tape reduction for n_param parameters (0: more
   parameters than it handles, OA_TAPE_NJ_MAX x 256) */
#define OA_TAPE_NJ_MAX 4
int tape_reduce_rows(int32_t n_param);
int adjoint_blocks_per_cu(int rng_kind);
/* no-allocation entry point: the device address of the per-GPU scratch array
   (call with `device` current), and the per-launch event setup kernel */
hipError_t noalloc_scratch(int device, void** ptr);
hipError_t noalloc_setup(int64_t* ev, int32_t* ev_id, unsigned long long* next, int64_t n_gs, int64_t n_photon,
                         int32_t event_id, hipStream_t s);
size_t adjoint_sort_bytes(int64_t n_hit, uint32_t n_det);
}
