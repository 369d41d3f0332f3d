/*
 * oa_merge.h : launchers of the hit-merge kernels (oa_merge.hip).
 */
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define OA_MERGE_PHOTON 0             /* sphoton records, 64 B */
#define OA_MERGE_PHOTONLITE 1         /* sphotonlite records, 16 B */
#define OA_MERGE_ALREADY_SELECTED 0xffffffffu

namespace oa_merge {

/* device scratch for up to n records (owned by the caller) */
struct Scratch {
    uint8_t* flags;       /* n */
    uint32_t* sel;        /* n selected indices, input order */
    uint32_t* vals;       /* n */
    uint32_t* vals2;      /* n sorted indices */
    uint64_t* keys;       /* n, reused for the unique keys */
    uint64_t* keys2;      /* n sorted keys */
    int32_t* run_len;     /* n */
    int32_t* run_off;     /* n */
    int32_t* counts;      /* 2: selected, runs */
    void* tmp;            /* hipCUB temporary storage */
    size_t tmp_bytes;
};

size_t scratch_bytes(int64_t n);
/* stable flagmask selection: indices into S.sel, count into S.counts[0] */
hipError_t select(const Scratch& S, int record, const void* d_in, int64_t n, uint32_t mask, hipStream_t s);
/* the selected records, input order, into d_out (n_cap slots launched, count read on the device) */
hipError_t gather_selected(const Scratch& S, int record, const void* d_in, void* d_out, int64_t n_cap, hipStream_t s);
/* keys of the n_sel selected records, stable sort, runs (count into S.counts[1]) */
hipError_t sort_runs(const Scratch& S, int record, const void* d_in, int64_t n_sel, float tw, hipStream_t s);
/* one merged record per run, in key order */
hipError_t reduce_runs(const Scratch& S, int record, const void* d_in, void* d_out, int64_t n_runs, hipStream_t s);

} // namespace oa_merge
