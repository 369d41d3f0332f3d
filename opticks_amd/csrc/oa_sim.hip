/*
 * oa_sim.hip : launchers of the propagation kernels for one thread-block size
 * (OA_SIM_BLOCK: 64, 128 or 256) and one random generator (OA_SIM_RNG:
 * 0 Philox, 1 precooked, 2 xorwow). The build compiles this file once per
 * pair, in parallel; a build with OA_SIM_ALL includes it for every pair into
 * oa_kernels.hip instead (one translation unit, as the OA_STATS counters need).
 */
#include "oa_sim.h"
#include "oa_sim_launch.h"

#if !defined(OA_SIM_BLOCK) || !defined(OA_SIM_RNG)
#error "compile with -DOA_SIM_BLOCK=<64|128|256> -DOA_SIM_RNG=<0|1|2>"
#endif

#if OA_SIM_RNG == 0
#define OA_SIM_R oa::PhiloxRng
#define OA_SIM_RT oa::PhiloxRngInline     /* the taped forward: the same stream, block inlined */
#elif OA_SIM_RNG == 1
#define OA_SIM_R oa::PrecookedRng
#define OA_SIM_RT OA_SIM_R
#else
#define OA_SIM_R oa::XorwowRng
#define OA_SIM_RT OA_SIM_R
#endif

namespace oa_launch {

hipError_t OA_SIM_NAME(sim_launch, OA_SIM_BLOCK, OA_SIM_RNG)(const SimParams& P, int kernel, int grid, hipStream_t s)
{
    constexpr int B = OA_SIM_BLOCK;
    const int f = (P.photonlite ? 1 : 0) | (P.ext_leaves ? 2 : 0);
    if (kernel == 0) {
        const unsigned nb = (unsigned)((P.n_photon + B - 1) / B);
        switch (f) {
            case 0: hipLaunchKernelGGL((k_sim_plain<B, 0, OA_SIM_R>), dim3(nb), dim3(B), 0, s, P); break;
            case 1: hipLaunchKernelGGL((k_sim_plain<B, 1, OA_SIM_R>), dim3(nb), dim3(B), 0, s, P); break;
            case 2: hipLaunchKernelGGL((k_sim_plain<B, 2, OA_SIM_R>), dim3(nb), dim3(B), 0, s, P); break;
            default: hipLaunchKernelGGL((k_sim_plain<B, 3, OA_SIM_R>), dim3(nb), dim3(B), 0, s, P); break;
        }
    } else if (P.tape) {
        /* a forward for a following adjoint: segment records on the tape (no lite output) */
        if (f & 2) hipLaunchKernelGGL((k_sim_wave<B, 2, OA_SIM_RT, oa::RecordTape>), dim3((unsigned)grid), dim3(B), 0, s, P);
        else hipLaunchKernelGGL((k_sim_wave<B, 0, OA_SIM_RT, oa::RecordTape>), dim3((unsigned)grid), dim3(B), 0, s, P);
    } else {
        switch (f) {
            case 0: hipLaunchKernelGGL((k_sim_wave<B, 0, OA_SIM_R>), dim3((unsigned)grid), dim3(B), 0, s, P); break;
            case 1: hipLaunchKernelGGL((k_sim_wave<B, 1, OA_SIM_R>), dim3((unsigned)grid), dim3(B), 0, s, P); break;
            case 2: hipLaunchKernelGGL((k_sim_wave<B, 2, OA_SIM_R>), dim3((unsigned)grid), dim3(B), 0, s, P); break;
            default: hipLaunchKernelGGL((k_sim_wave<B, 3, OA_SIM_R>), dim3((unsigned)grid), dim3(B), 0, s, P); break;
        }
    }
    return hipGetLastError();
}

int OA_SIM_NAME(sim_occupancy, OA_SIM_BLOCK, OA_SIM_RNG)()
{
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_sim_wave<OA_SIM_BLOCK, 0, OA_SIM_R>, OA_SIM_BLOCK, 0) !=
        hipSuccess)
        return 0;
    return nb;
}

#if OA_SIM_BLOCK == OA_ADJ_BLOCK
hipError_t OA_SIM_NAME(adjoint_launch, OA_ADJ_BLOCK, OA_SIM_RNG)(const SimParams& P, const AdjParams& A,
                                                                 const float4* hits, int64_t begin, int64_t n_max,
                                                                 const int64_t* hit_off, int32_t n_ev, int64_t hit_cap,
                                                                 float* S, int grid, hipStream_t s)
{
    if (P.ext_leaves)
        hipLaunchKernelGGL((k_adjoint_wave<OA_ADJ_BLOCK, true, OA_SIM_R>), dim3((unsigned)grid), dim3(OA_ADJ_BLOCK), 0,
                           s, P, A, hits, begin, n_max, hit_off, n_ev, hit_cap, S);
    else
        hipLaunchKernelGGL((k_adjoint_wave<OA_ADJ_BLOCK, false, OA_SIM_R>), dim3((unsigned)grid), dim3(OA_ADJ_BLOCK), 0,
                           s, P, A, hits, begin, n_max, hit_off, n_ev, hit_cap, S);
    return hipGetLastError();
}

int OA_SIM_NAME(adjoint_occupancy, OA_ADJ_BLOCK, OA_SIM_RNG)()
{
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_adjoint_wave<OA_ADJ_BLOCK, false, OA_SIM_R>, OA_ADJ_BLOCK,
                                                     0) != hipSuccess)
        return 0;
    return nb;
}
#endif

} // namespace oa_launch

#undef OA_SIM_R
#undef OA_SIM_RT
