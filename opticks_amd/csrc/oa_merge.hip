/*
 * oa_merge.hip : hit selection and merging by (sensor identity, time bucket).
 *
 * Restates SPM::merge_partial_select (sysrap/SPM.cu:135-240) for MI355X:
 *   0. select records with flagmask & mask != 0 (skipped for the "already
 *      selected" mask), stable;
 *   1. time window 0: return the selection in input order;
 *   2-4. key = (identity << 48) | (unsigned)(time / window) as a 64-bit
 *      integer, then a stable sort of the selected records by key; here the
 *      same order is sorted as (identity & 0xffff) << 32 | bucket over its 48
 *      significant bits (the shift by 48 keeps the same 16 identity bits);
 *   6. reduce each run of equal keys with the record type's reduce_op.
 * Record types and their functors:
 *   sphoton     64 B  key/reduce/select  sysrap/sphoton.h:292-320
 *   sphotonlite 16 B  key/reduce/select  sysrap/sphotonlite.h:68-97
 * The thrust calls become hipCUB device-wide primitives (flagged select,
 * radix sort of (key, index) pairs, run-length encode, exclusive scan) and a
 * run-reduction kernel that folds each run left to right in sorted order, so
 * ties in time keep the earliest record in input order.
 */
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>
#include <stdint.h>
#include "oa_merge.h"

namespace {

struct PhotonAcc {
    /* sphoton: q0 (pos, time), q1 (mom, hitcount_iindex), q2 (pol, wavelength),
       q3 (orient_boundary_flag, identity, index, flagmask) */
    static constexpr int WORDS = 16;
    __device__ static uint32_t identity(const uint32_t* r) { return r[13] & 0x00ffffffu; }
    __device__ static float time(const uint32_t* r) { return __uint_as_float(r[3]); }
    __device__ static uint32_t flagmask(const uint32_t* r) { return r[15]; }
};

struct LiteAcc {
    /* sphotonlite: hitcount_identity (hi16 hitcount, lo16 identity), time, lpos, flagmask */
    static constexpr int WORDS = 4;
    __device__ static uint32_t identity(const uint32_t* r) { return r[0] & 0x0000ffffu; }
    __device__ static float time(const uint32_t* r) { return __uint_as_float(r[1]); }
    __device__ static uint32_t flagmask(const uint32_t* r) { return r[3]; }
};

__global__ void k_flags(const uint32_t* in, int words, int flag_word, int64_t n, uint32_t mask, uint8_t* flags)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    flags[i] = (mask == OA_MERGE_ALREADY_SELECTED || (in[i * words + flag_word] & mask) != 0u) ? 1 : 0;
}

template <class Acc>
__global__ void k_keys(const uint32_t* in, const uint32_t* sel, const int32_t* d_nsel, float tw, uint64_t* keys,
                       uint32_t* vals)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= *d_nsel) return;
    const uint32_t* r = in + (int64_t)sel[i] * Acc::WORDS;
    const uint32_t id = Acc::identity(r) & 0xffffu;           /* 16 bits of identity in the key */
    const uint32_t bucket = (uint32_t)(Acc::time(r) / tw);
    keys[i] = ((uint64_t)id << 32) | (uint64_t)bucket;          /* 48 significant bits */
    vals[i] = sel[i];
}

__global__ void k_gather(const uint32_t* in, int words, const uint32_t* idx, const int32_t* d_n, uint32_t* out)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t i = t / words, w = t % words;
    if (i >= *d_n) return;
    out[i * words + w] = in[(int64_t)idx[i] * words + w];
}

/* sphoton::reduce_op folded left to right over one run:
   r = (fminf(a.time, b.time) == a.time) ? a : b; r.flagmask = a.flagmask | b.flagmask;
   r.hitcount = a.hitcount + b.hitcount (16-bit field) */
__global__ void k_reduce_photon(const uint32_t* in, const uint32_t* order, const int32_t* run_off,
                                const int32_t* run_len, const int32_t* d_nruns, uint32_t* out)
{
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= *d_nruns) return;
    const int32_t o = run_off[k], c = run_len[k];
    uint32_t best = order[o];
    float bt = __uint_as_float(in[(int64_t)best * 16 + 3]);
    uint32_t fm = in[(int64_t)best * 16 + 15];
    uint32_t hc = in[(int64_t)best * 16 + 7] >> 16;
    int32_t j = 1;
    /* four records' words in flight per step, folded in the same order as one at a time */
    for (; j + 4 <= c; j += 4) {
        uint32_t b[4], tw[4], fw[4], hw[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) b[u] = order[o + j + u];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            tw[u] = in[(int64_t)b[u] * 16 + 3];
            fw[u] = in[(int64_t)b[u] * 16 + 15];
            hw[u] = in[(int64_t)b[u] * 16 + 7];
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const float t = __uint_as_float(tw[u]);
            if (!(fminf(bt, t) == bt)) { best = b[u]; bt = t; }
            fm |= fw[u];
            hc = (hc + (hw[u] >> 16)) & 0xffffu;
        }
    }
    for (; j < c; ++j) {
        const uint32_t b = order[o + j];
        const float t = __uint_as_float(in[(int64_t)b * 16 + 3]);
        if (!(fminf(bt, t) == bt)) { best = b; bt = t; }
        fm |= in[(int64_t)b * 16 + 15];
        hc = (hc + (in[(int64_t)b * 16 + 7] >> 16)) & 0xffffu;
    }
    const uint32_t* r = in + (int64_t)best * 16;
    uint32_t* w = out + k * 16;
    for (int q = 0; q < 16; ++q) w[q] = r[q];
    w[7] = (r[7] & 0x0000ffffu) | (hc << 16);
    w[15] = fm;
}

/* sphotonlite::reduce_op folded left to right: r = a; r.time = fminf(a.time, b.time);
   r.flagmask |= b.flagmask; r.hitcount = a.hitcount + b.hitcount, identity of a */
__global__ void k_reduce_lite(const uint32_t* in, const uint32_t* order, const int32_t* run_off,
                              const int32_t* run_len, const int32_t* d_nruns, uint32_t* out)
{
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= *d_nruns) return;
    const int32_t o = run_off[k], c = run_len[k];
    const uint32_t* a = in + (int64_t)order[o] * 4;
    uint32_t hci = a[0];
    float t = __uint_as_float(a[1]);
    uint32_t fm = a[3];
    int32_t j = 1;
    /* four records in flight per step, folded in the same order as one at a time */
    for (; j + 4 <= c; j += 4) {
        const uint32_t* b[4];
        uint32_t hw[4], tw[4], fw[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) b[u] = in + (int64_t)order[o + j + u] * 4;
#pragma unroll
        for (int u = 0; u < 4; ++u) { hw[u] = b[u][0]; tw[u] = b[u][1]; fw[u] = b[u][3]; }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            t = fminf(t, __uint_as_float(tw[u]));
            fm |= fw[u];
            const uint32_t hc = ((hci >> 16) + (hw[u] >> 16)) & 0xffffu;
            hci = (hc << 16) | (hci & 0x0000ffffu);
        }
    }
    for (; j < c; ++j) {
        const uint32_t* b = in + (int64_t)order[o + j] * 4;
        t = fminf(t, __uint_as_float(b[1]));
        fm |= b[3];
        const uint32_t hc = ((hci >> 16) + (b[0] >> 16)) & 0xffffu;
        hci = (hc << 16) | (hci & 0x0000ffffu);
    }
    uint32_t* w = out + k * 4;
    w[0] = hci;
    w[1] = __float_as_uint(t);
    w[2] = a[2];
    w[3] = fm;
}

inline unsigned nblk(int64_t n, int b) { return (unsigned)((n + b - 1) / b); }

} // namespace

namespace oa_merge {

/* hipCUB temporary storage needed for n records */
size_t scratch_bytes(int64_t n)
{
    size_t a = 0, b = 0, c = 0, d = 0;
    const int ni = (int)n;
    (void)hipcub::DeviceSelect::Flagged(nullptr, a, hipcub::CountingInputIterator<uint32_t>(0), (uint8_t*)nullptr,
                                        (uint32_t*)nullptr, (int32_t*)nullptr, ni);
    (void)hipcub::DeviceRadixSort::SortPairs(nullptr, b, (uint64_t*)nullptr, (uint64_t*)nullptr, (uint32_t*)nullptr,
                                             (uint32_t*)nullptr, ni);
    (void)hipcub::DeviceRunLengthEncode::Encode(nullptr, c, (uint64_t*)nullptr, (uint64_t*)nullptr, (int32_t*)nullptr,
                                                (int32_t*)nullptr, ni);
    (void)hipcub::DeviceScan::ExclusiveSum(nullptr, d, (int32_t*)nullptr, (int32_t*)nullptr, ni);
    return std::max(std::max(a, b), std::max(c, d));
}

hipError_t select(const Scratch& S, int record, const void* d_in, int64_t n, uint32_t mask, hipStream_t s)
{
    const int words = record == OA_MERGE_PHOTON ? 16 : 4;
    const int flag_word = record == OA_MERGE_PHOTON ? 15 : 3;
    hipLaunchKernelGGL(k_flags, dim3(nblk(n, 256)), dim3(256), 0, s, (const uint32_t*)d_in, words, flag_word, n, mask,
                       S.flags);
    size_t bytes = S.tmp_bytes;
    return hipcub::DeviceSelect::Flagged(S.tmp, bytes, hipcub::CountingInputIterator<uint32_t>(0), S.flags, S.sel,
                                         S.counts, (int)n, s);
}

hipError_t gather_selected(const Scratch& S, int record, const void* d_in, void* d_out, int64_t n_cap, hipStream_t s)
{
    const int words = record == OA_MERGE_PHOTON ? 16 : 4;
    hipLaunchKernelGGL(k_gather, dim3(nblk(n_cap * words, 256)), dim3(256), 0, s, (const uint32_t*)d_in, words, S.sel,
                       S.counts, (uint32_t*)d_out);
    return hipGetLastError();
}

hipError_t sort_runs(const Scratch& S, int record, const void* d_in, int64_t n_sel, float tw, hipStream_t s)
{
    if (n_sel <= 0) return hipSuccess;
    if (record == OA_MERGE_PHOTON)
        hipLaunchKernelGGL(k_keys<PhotonAcc>, dim3(nblk(n_sel, 256)), dim3(256), 0, s, (const uint32_t*)d_in, S.sel,
                           S.counts, tw, S.keys, S.vals);
    else
        hipLaunchKernelGGL(k_keys<LiteAcc>, dim3(nblk(n_sel, 256)), dim3(256), 0, s, (const uint32_t*)d_in, S.sel,
                           S.counts, tw, S.keys, S.vals);
    size_t bytes = S.tmp_bytes;
    hipError_t e = hipcub::DeviceRadixSort::SortPairs(S.tmp, bytes, S.keys, S.keys2, S.vals, S.vals2, (int)n_sel, 0, 48, s);
    if (e != hipSuccess) return e;
    bytes = S.tmp_bytes;
    e = hipcub::DeviceRunLengthEncode::Encode(S.tmp, bytes, S.keys2, S.keys, S.run_len, S.counts + 1, (int)n_sel, s);
    if (e != hipSuccess) return e;
    bytes = S.tmp_bytes;
    return hipcub::DeviceScan::ExclusiveSum(S.tmp, bytes, S.run_len, S.run_off, (int)n_sel, s);
}

hipError_t reduce_runs(const Scratch& S, int record, const void* d_in, void* d_out, int64_t n_runs, hipStream_t s)
{
    if (n_runs <= 0) return hipSuccess;
    if (record == OA_MERGE_PHOTON)
        hipLaunchKernelGGL(k_reduce_photon, dim3(nblk(n_runs, 256)), dim3(256), 0, s, (const uint32_t*)d_in, S.vals2,
                           S.run_off, S.run_len, S.counts + 1, (uint32_t*)d_out);
    else
        hipLaunchKernelGGL(k_reduce_lite, dim3(nblk(n_runs, 256)), dim3(256), 0, s, (const uint32_t*)d_in, S.vals2,
                           S.run_off, S.run_len, S.counts + 1, (uint32_t*)d_out);
    return hipGetLastError();
}

} // namespace oa_merge
