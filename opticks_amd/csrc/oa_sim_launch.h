/* oa_sim_launch.h : launchers of the propagation kernels, one per
   (block size, random generator) pair, defined in oa_sim.hip. */
#pragma once
#include <hip/hip_runtime.h>
#include "oa_layout.h"

struct AdjParams;

#define OA_SIM_CAT3(a, b, c) a##_##b##_##c
#define OA_SIM_NAME(a, b, c) OA_SIM_CAT3(a, b, c)

#define OA_SIM_DECLARE(B, K)                                                                   \
    hipError_t sim_launch_##B##_##K(const SimParams& P, int kernel, int grid, hipStream_t s); \
    int sim_occupancy_##B##_##K();

namespace oa_launch {
OA_SIM_DECLARE(64, 0)
OA_SIM_DECLARE(64, 1)
OA_SIM_DECLARE(64, 2)
OA_SIM_DECLARE(128, 0)
OA_SIM_DECLARE(128, 1)
OA_SIM_DECLARE(128, 2)
OA_SIM_DECLARE(256, 0)
OA_SIM_DECLARE(256, 1)
OA_SIM_DECLARE(256, 2)

#define OA_ADJ_DECLARE(K)                                                                                     \
    hipError_t adjoint_launch_128_##K(const SimParams& P, const AdjParams& A, const float4* hits, int64_t begin, \
                                      int64_t n_max, const int64_t* hit_off, int32_t n_ev, int64_t hit_cap,    \
                                      float* S, int grid, hipStream_t s);                                      \
    int adjoint_occupancy_128_##K();
OA_ADJ_DECLARE(0)
OA_ADJ_DECLARE(1)
OA_ADJ_DECLARE(2)
}
