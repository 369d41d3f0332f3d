// ref_csg.cc : exports the reference's own CPU-compilable CSG intersection
// (CSG/csg_intersect_leaf.h, csg_intersect_node.h, csg_intersect_tree.h with
// the no-CUDA vector shim sysrap/scuda.h -> snocuda_vector_functions.h) as a C
// symbol, so tests can compare the oracle's restatement with the reference code
// itself. TEST INFRASTRUCTURE ONLY: built by oracle/Makefile into oracle/_ref/
// from the sources where they lie under the reference checkout; no reference
// source is copied into this repository.
#include "scuda.h"
#include "squad.h"
#include "sqat4.h"
#include "csg_intersect_leaf.h"
#include "csg_intersect_node.h"
#include "csg_intersect_tree.h"

extern "C" int ref_intersect_prim(const float* root_node, const float* itra, const float* plan,
                                  float t_min, const float* o, const float* d, float* isect4)
{
    float4 isect = make_float4(0.f, 0.f, 0.f, 0.f);
    float3 ro = make_float3(o[0], o[1], o[2]);
    float3 rd = make_float3(d[0], d[1], d[2]);
    bool valid = intersect_prim(isect, (const CSGNode*)root_node, (const float4*)plan,
                                (const qat4*)itra, t_min, ro, rd, false);
    isect4[0] = isect.x; isect4[1] = isect.y; isect4[2] = isect.z; isect4[3] = isect.w;
    return valid ? 1 : 0;
}

extern "C" int ref_lut_lookup(int op, int stateA, int stateB, int acloser)
{
    LUT lut;
    return lut.lookup((OpticksCSG_t)op, (IntersectionState_t)stateA, (IntersectionState_t)stateB, acloser != 0);
}
