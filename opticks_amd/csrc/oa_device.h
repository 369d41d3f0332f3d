/*
 * oa_device.h : device-side optical-photon propagation for gfx950 (CDNA4).
 *
 * One lane carries one photon through generation and the bounce loop of
 * CSGOptiX/CSGOptiX7.cu:432-530. The OptiX trace is replaced by a two-level
 * BVH (TLAS over instances, BLAS over the CSGPrim of each solid) traversed
 * with a per-lane short stack in LDS; the intersection program is the CSG
 * single-hit algorithm of CSG/csg_intersect_tree.h; the physics follows
 * qudarap/qsim.h with boundary tables read as FP32-lerped rows in HBM
 * instead of hardware textures.
 *
 * Floating point: compiled with -ffp-contract=off and written in the
 * reference's operand order, so that results are bit-identical to the CPU
 * checker on the same random streams. sqrtf and '/' are IEEE correctly
 * rounded under hipcc's defaults; log/sin/cos come from oa_portable.h.
 */
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include "oa_portable.h"
#include "oa_layout.h"

#define OA_D __device__ static inline __attribute__((always_inline))

#ifdef OA_STATS
__device__ unsigned long long oa_stats[OA_ST_N];
#define OA_COUNT(i) atomicAdd(&oa_stats[(i)], 1ull)
#define OA_COUNT_N(i, n) atomicAdd(&oa_stats[(i)], (unsigned long long)(n))
/* cycles between a mark and an add, both at points where the wave's active lanes are
   converged; counted once per wave */
#define OA_CYC_MARK(v) const unsigned long long v = (unsigned long long)__builtin_readcyclecounter()
#define OA_CYC_ADD(i, v)                                                                      \
    do {                                                                                      \
        const unsigned long long t_ = (unsigned long long)__builtin_readcyclecounter();       \
        const unsigned long long m_ = __ballot(1);                                            \
        if ((int)__lane_id() == __ffsll((long long)m_) - 1) atomicAdd(&oa_stats[(i)], t_ - (v)); \
    } while (0)
#define OA_WAVE_COUNT(w, l)                                                                   \
    do {                                                                                      \
        const unsigned long long m_ = __ballot(1);                                            \
        if ((int)__lane_id() == __ffsll((long long)m_) - 1) {                                 \
            atomicAdd(&oa_stats[(w)], 1ull);                                                  \
            atomicAdd(&oa_stats[(l)], (unsigned long long)__popcll(m_));                      \
        }                                                                                     \
    } while (0)
/* per-trace iteration histograms (log2 bins) of the TLAS walk [0] and the
   leaf walk [1], and the first OA_SLOW_MAX traces whose TLAS walk took more
   than OA_SLOW_ITER iterations: (o, tmin), (d, tmax), (best_t, n1, n2, inst) */
#define OA_SLOW_MAX 16384
#define OA_SLOW_ITER 100
__device__ unsigned long long oa_hist[2][32];
__device__ unsigned long long oa_slow_n;
__device__ float4 oa_slow[OA_SLOW_MAX * 3];
#else
#define OA_WAVE_COUNT(w, l) ((void)0)
#define OA_CYC_MARK(v) ((void)0)
#define OA_CYC_ADD(i, v) ((void)0)
#define OA_COUNT(i) ((void)0)
#define OA_COUNT_N(i, n) ((void)0)
#endif

namespace oa {

/* ----------------------------------------------------------- vectors */

struct v3 { float x, y, z; };

OA_D v3 mk(float x, float y, float z) { v3 r; r.x = x; r.y = y; r.z = z; return r; }
OA_D v3 operator+(v3 a, v3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
OA_D v3 operator-(v3 a, v3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
OA_D v3 operator-(v3 a) { return mk(-a.x, -a.y, -a.z); }
OA_D v3 operator*(float s, v3 a) { return mk(a.x * s, a.y * s, a.z * s); }
OA_D v3 operator*(v3 a, float s) { return mk(a.x * s, a.y * s, a.z * s); }
OA_D float dot(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
OA_D v3 cross(v3 a, v3 b) { return mk(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x); }
OA_D float length(v3 v) { return sqrtf(dot(v, v)); }
OA_D v3 normalize(v3 v) { float inv = 1.0f / sqrtf(dot(v, v)); return v * inv; }
OA_D uint32_t fbits(float f) { return __float_as_uint(f); }
OA_D bool sgnbit(float f) { return (__float_as_uint(f) >> 31) != 0u; }

#define OA_PI_F 3.14159274101257324f

/* --------------------------------------------------------- constants */

enum : uint32_t {
    F_CERENKOV = 0x1u << 0, F_SCINTILLATION = 0x1u << 1, F_TORCH = 0x1u << 2,
    F_BULK_ABSORB = 0x1u << 3, F_BULK_REEMIT = 0x1u << 4, F_BULK_SCATTER = 0x1u << 5,
    F_SURFACE_DETECT = 0x1u << 6, F_SURFACE_ABSORB = 0x1u << 7,
    F_SURFACE_DREFLECT = 0x1u << 8, F_SURFACE_SREFLECT = 0x1u << 9,
    F_BOUNDARY_REFLECT = 0x1u << 10, F_BOUNDARY_TRANSMIT = 0x1u << 11
};
enum : int { CMD_BREAK = 1, CMD_CONTINUE = 2, CMD_BOUNDARY = 3 };
enum : int { EMS_NOSURFACE = 1, EMS_SURFACE = 2, EMS_ZPLUS_SENSOR_A = 3 };
enum : uint32_t { TC_ZERO = 0, TC_UNION = 1, TC_INTERSECTION = 2, TC_DIFFERENCE = 3, TC_NODE = 11,
                  TC_CONTIGUOUS = 11, TC_DISCONTIGUOUS = 12, TC_OVERLAP = 13,
                  TC_LEAF = 101, TC_SPHERE = 101, TC_ZSPHERE = 103, TC_CYLINDER = 105, TC_CONE = 108,
                  TC_BOX3 = 110, TC_CONVEXPOLYHEDRON = 112, TC_HYPERBOLOID = 117, TC_PHICUT = 121,
                  TC_HALFSPACE = 125 };

/* ------------------------------------------------------------ random */

/* The per-photon generator is a template parameter of the propagation code:
   each kernel instantiation carries only the state of its own kind. */
enum : int { RNG_PHILOX = 0, RNG_PRECOOKED = 1, RNG_XORWOW = 2 };

struct PhiloxRng { oa_philox_stream ph; };
/* the same stream with its block generated inline at each draw site (the taped
   forward's instantiation, whose register allocation measured better with it) */
struct PhiloxRngInline { oa_philox_stream ph; };
/* whether the physics of a kernel drawing from R calls the table-index wrap's integer
   divisions out of line (a smaller kernel), or inline (the taped forward) */
template <class R> struct wrap_ool { static constexpr bool value = true; };
template <> struct wrap_ool<PhiloxRngInline> { static constexpr bool value = false; };
struct XorwowRng { oa_xorwow_state xw; };
/* precooked uniforms (the reference's s_seq aligned-randoms mode): a row of nk
   values per photon; the cursor wraps modulo nk as s_seq::flat recycles
   (sysrap/s_seq.h:105-129) */
struct PrecookedRng {
    const float* row;
    int32_t cursor, nk;
};

/* the Philox block out of line: one copy of its ten rounds instead of one per draw site */
__device__ static __attribute__((noinline)) uint4 philox_block(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3,
                                                                uint32_t k0, uint32_t k1)
{
    oa_u32x4 c;
    c.v[0] = c0; c.v[1] = c1; c.v[2] = c2; c.v[3] = c3;
    const oa_u32x4 o = oa_philox4x32_10(c, k0, k1);
    return make_uint4(o.v[0], o.v[1], o.v[2], o.v[3]);
}
/* oa_rng_uniform with the block generated out of line (the same words) */
OA_D float uni(PhiloxRng& r)
{
    oa_philox_stream* s = &r.ph;
    const uint32_t x = s->b0;
    s->b0 = s->b1; s->b1 = s->b2; s->b2 = s->b3;
    s->pos += 1u;
    if ((s->pos & 3u) == 0u) {
        const uint64_t block = s->pos >> 2;
        const uint4 o = philox_block((uint32_t)block, (uint32_t)(block >> 32), (uint32_t)s->photon,
                                     (uint32_t)(s->photon >> 32), (uint32_t)s->key, (uint32_t)(s->key >> 32));
        s->b0 = o.x; s->b1 = o.y; s->b2 = o.z; s->b3 = o.w;
    }
    return oa_u32_to_uniform(x);
}
OA_D float uni(PhiloxRngInline& r) { return oa_rng_uniform(&r.ph); }
OA_D float uni(XorwowRng& r) { return oa_xorwow_uniform(&r.xw); }
OA_D float uni(PrecookedRng& r)
{
    const float v = r.row[r.cursor];
    r.cursor = r.cursor + 1 == r.nk ? 0 : r.cursor + 1;
    return v;
}

/* Rejection loops of the reference (marsaglia, Rayleigh, torch sphere,
   Cerenkov position) are unbounded; here each stops after OA_MAX_TRIALS
   trials and keeps its last trial, so no input (degenerate gensteps, a
   precooked row that never accepts) can hold a lane forever. The acceptance
   probabilities are pi/4 (marsaglia, full-range torch sphere), 2/3 on
   average (Rayleigh) and >= 1/2 (Cerenkov, positive yields): a random stream
   reaches the bound with probability below (1/2)^65536. */
#define OA_MAX_TRIALS 65536

/* ------------------------------------------------------------ photon */

struct Photon {
    v3 pos; float time;
    v3 mom; uint32_t hitcount_iindex;
    v3 pol; float wavelength;
    uint32_t obf, identity, index, flagmask;   /* orient_boundary_flag ... (sysrap/sphoton.h:171-246) */
};

OA_D uint32_t ph_flag(const Photon& p) { return p.obf & 0xffffu; }
OA_D uint32_t ph_boundary(const Photon& p) { return (p.obf & 0x7fff0000u) >> 16; }
OA_D void set_flag(Photon& p, uint32_t flag)
{
    p.obf = (p.obf & 0xffff0000u) | (flag & 0xffffu);
    p.flagmask |= flag;
}

/* per-trace record, the quad2 PRD of sysrap/squad.h:180-250 */
struct Prd {
    v3 normal; float t;
    float lposcost;
    float lx, ly;                 /* local hit position x, y (azimuth for sphotonlite); miss: 0, NaN */
    uint32_t gprim;               /* globalPrimIdx of the hit prim (simtrace records) */
    uint32_t iindex_identity, boundary;
};

/* ------------------------------------------- robust quadratic roots */
/* CSG/csg_robust_quadratic_roots.h:104-136 */

OA_D void rq_roots(float& t1, float& t2, float& sdisc, float d, float b, float c)
{
    float disc = b * b - d * c;
    sdisc = disc > 0.f ? sqrtf(disc) : 0.f;
    float q = b > 0.f ? -(b + sdisc) : -(b - sdisc);
    float root1 = q / d;
    float root2 = c / q;
    t1 = fminf(root1, root2);
    t2 = fmaxf(root1, root2);
}

OA_D void rq_roots_disq(float t_min, float& t1, float& t2, float& sdisc, float d, float b, float c)
{
    float disc = b * b - d * c;
    sdisc = disc > 0.f ? sqrtf(disc) : 0.f;
    float q = b > 0.f ? -(b + sdisc) : -(b - sdisc);
    float root1 = sdisc > 0.f ? q / d : t_min;
    float root2 = sdisc > 0.f ? c / q : t_min;
    t1 = fminf(root1, root2);
    t2 = fmaxf(root1, root2);
}

/* -------------------------------------------------------------- leaves */

struct Isect { float x, y, z, w; };

/* value select: keeps Isect arrays in registers (a select of two addresses
   would pin them to scratch memory) */
OA_D Isect pick(bool c, const Isect& a, const Isect& b)
{
    Isect r;
    r.x = c ? a.x : b.x; r.y = c ? a.y : b.y; r.z = c ? a.z : b.z; r.w = c ? a.w : b.w;
    return r;
}

OA_D bool leaf_sphere(Isect& is, float4 q0, float t_min, v3 ro, v3 rd)
{
    v3 O = ro - mk(q0.x, q0.y, q0.z);
    float radius = q0.w;
    float b = dot(O, rd);
    float c = dot(O, O) - radius * radius;
    float d = dot(rd, rd);
    float r1, r2, sdisc;
    rq_roots(r1, r2, sdisc, d, b, c);
    float t_cand = sdisc > 0.f ? (r1 > t_min ? r1 : r2) : t_min;
    bool valid = t_cand > t_min;
    if (valid) {
        is.x = (O.x + t_cand * rd.x) / radius;
        is.y = (O.y + t_cand * rd.y) / radius;
        is.z = (O.z + t_cand * rd.z) / radius;
        is.w = t_cand;
    }
    return valid;
}

OA_D bool leaf_zsphere(Isect& is, float4 q0, float4 q1, float t_min, v3 ro, v3 rd)
{
    v3 O = ro - mk(q0.x, q0.y, q0.z);
    float radius = q0.w;
    float b = dot(O, rd);
    float c = dot(O, O) - radius * radius;
    if (c > 0.f && b > 0.f) return false;
    float zmax = q0.z + q1.y;
    float zmin = q0.z + q1.x;
    float d = dot(rd, rd);
    float t1sph, t2sph, sdisc;
    rq_roots(t1sph, t2sph, sdisc, d, b, c);
    float z1sph = ro.z + t1sph * rd.z;
    float z2sph = ro.z + t2sph * rd.z;
    float idz = 1.f / rd.z;
    float t_QCAP = (zmax - ro.z) * idz;
    float t_PCAP = (zmin - ro.z) * idz;
    float t1cap = fminf(t_QCAP, t_PCAP);
    float t2cap = fmaxf(t_QCAP, t_PCAP);
    if (t1cap < t1sph || t1cap > t2sph) t1cap = t_min;
    if (t2cap < t1sph || t2cap > t2sph) t2cap = t_min;
    float t_cand = t_min;
    if (sdisc > 0.f) {
        if (t1sph > t_min && z1sph > zmin && z1sph <= zmax) t_cand = t1sph;
        else if (t1cap > t_min) t_cand = t1cap;
        else if (t2cap > t_min) t_cand = t2cap;
        else if (t2sph > t_min && z2sph > zmin && z2sph <= zmax) t_cand = t2sph;
    }
    bool valid = t_cand > t_min;
    if (valid) {
        is.w = t_cand;
        if (t_cand == t1sph || t_cand == t2sph) {
            is.x = (O.x + t_cand * rd.x) / radius;
            is.y = (O.y + t_cand * rd.y) / radius;
            is.z = (O.z + t_cand * rd.z) / radius;
        } else {
            is.x = 0.f;
            is.y = 0.f;
            is.z = t_cand == t_PCAP ? -1.f : 1.f;
        }
    }
    return valid;
}

OA_D bool leaf_cylinder(Isect& is, float4 q0, float4 q1, float t_min, v3 ro, v3 rd)
{
    const float r = q0.w, z1 = q1.x, z2 = q1.y;
    const float ox = ro.x, oy = ro.y, oz = ro.z, vx = rd.x, vy = rd.y, vz = rd.z;
    const float r2 = r * r;
    const float a = vx * vx + vy * vy;
    const float b = ox * vx + oy * vy;
    const float c = ox * ox + oy * oy - r2;
    float t_near, t_far, sdisc;
    rq_roots_disq(t_min, t_near, t_far, sdisc, a, b, c);
    float z_near = oz + t_near * vz;
    float z_far = oz + t_far * vz;
    const float t_z1cap = (z1 - oz) / vz;
    const float r2_z1cap = (ox + t_z1cap * vx) * (ox + t_z1cap * vx) + (oy + t_z1cap * vy) * (oy + t_z1cap * vy);
    const float t_z2cap = (z2 - oz) / vz;
    const float r2_z2cap = (ox + t_z2cap * vx) * (ox + t_z2cap * vx) + (oy + t_z2cap * vy) * (oy + t_z2cap * vy);
    float t_cand = __builtin_inff();
    if (t_near > t_min && z_near > z1 && z_near < z2 && t_near < t_cand) t_cand = t_near;
    if (t_far > t_min && z_far > z1 && z_far < z2 && t_far < t_cand) t_cand = t_far;
    if (t_z1cap > t_min && r2_z1cap <= r2 && t_z1cap < t_cand) t_cand = t_z1cap;
    if (t_z2cap > t_min && r2_z2cap <= r2 && t_z2cap < t_cand) t_cand = t_z2cap;
    bool valid = t_cand > t_min && t_cand < __builtin_inff();
    if (valid) {
        bool sheet = (t_cand == t_near || t_cand == t_far);
        is.x = sheet ? (ox + t_cand * vx) / r : 0.f;
        is.y = sheet ? (oy + t_cand * vy) / r : 0.f;
        is.z = sheet ? 0.f : (t_cand == t_z1cap ? -1.f : 1.f);
        is.w = t_cand;
    }
    return valid;
}

OA_D bool leaf_box3(Isect& is, float4 q0, float t_min, v3 ro, v3 rd)
{
    const v3 bmin = mk(-q0.x / 2.f, -q0.y / 2.f, -q0.z / 2.f);
    const v3 bmax = mk(q0.x / 2.f, q0.y / 2.f, q0.z / 2.f);
    const v3 idir = mk(1.f / rd.x, 1.f / rd.y, 1.f / rd.z);
    v3 t0 = mk((bmin.x - ro.x) * idir.x, (bmin.y - ro.y) * idir.y, (bmin.z - ro.z) * idir.z);
    v3 t1 = mk((bmax.x - ro.x) * idir.x, (bmax.y - ro.y) * idir.y, (bmax.z - ro.z) * idir.z);
    float t_near = fmaxf(fmaxf(fminf(t0.x, t1.x), fminf(t0.y, t1.y)), fminf(t0.z, t1.z));
    float t_far = fminf(fminf(fmaxf(t0.x, t1.x), fmaxf(t0.y, t1.y)), fmaxf(t0.z, t1.z));
    bool along_x = rd.x != 0.f && rd.y == 0.f && rd.z == 0.f;
    bool along_y = rd.x == 0.f && rd.y != 0.f && rd.z == 0.f;
    bool along_z = rd.x == 0.f && rd.y == 0.f && rd.z != 0.f;
    bool in_x = ro.x > bmin.x && ro.x < bmax.x;
    bool in_y = ro.y > bmin.y && ro.y < bmax.y;
    bool in_z = ro.z > bmin.z && ro.z < bmax.z;
    bool has;
    if (along_x) has = in_y && in_z;
    else if (along_y) has = in_x && in_z;
    else if (along_z) has = in_x && in_y;
    else has = (t_far > t_near && t_far > 0.f);
    if (!has) return false;
    float t_cand = t_min < t_near ? t_near : (t_min < t_far ? t_far : t_min);
    v3 p = (ro + t_cand * rd) - mk(0.f, 0.f, 0.f);
    v3 pa = mk(fabsf(p.x) / (bmax.x - bmin.x), fabsf(p.y) / (bmax.y - bmin.y), fabsf(p.z) / (bmax.z - bmin.z));
    v3 n = mk(0.f, 0.f, 0.f);
    if (pa.x >= pa.y && pa.x >= pa.z) n.x = copysignf(1.f, p.x);
    else if (pa.y >= pa.x && pa.y >= pa.z) n.y = copysignf(1.f, p.y);
    else if (pa.z >= pa.x && pa.z >= pa.y) n.z = copysignf(1.f, p.z);
    if (!(t_cand > t_min)) return false;
    is.x = n.x; is.y = n.y; is.z = n.z; is.w = t_cand;
    return true;
}

/* qat4 right_multiply / left_multiply_inplace (sysrap/sqat4.h:45-104) on a node transform */
OA_D v3 q_right(const float4* q, v3 v, float w)
{
    float4 a = q[0], b = q[1], c = q[2], d = q[3];
    return mk(a.x * v.x + b.x * v.y + c.x * v.z + d.x * w,
              a.y * v.x + b.y * v.y + c.y * v.z + d.y * w,
              a.z * v.x + b.z * v.y + c.z * v.z + d.z * w);
}
OA_D void q_left_inplace(const float4* q, Isect& v, float w)
{
    float4 a = q[0], b = q[1], c = q[2];
    float x = a.x * v.x + a.y * v.y + a.z * v.z + a.w * w;
    float y = b.x * v.x + b.y * v.y + b.z * v.z + b.w * w;
    float z = c.x * v.x + c.y * v.y + c.z * v.z + c.w * w;
    v.x = x; v.y = y; v.z = z;
}

#define OA_RT_DEFAULT_MAX 1.e27f
OA_D float fmin4(float a, float b, float c, float d) { return fminf(fminf(a, b), fminf(c, d)); }

/* truncated cone with caps (CSG/csg_intersect_leaf_newcone.h:45-123) */
OA_D bool leaf_cone(Isect& is, float4 q0, float t_min, v3 o, v3 d)
{
    const float r1 = q0.x, z1 = q0.y, r2 = q0.z, z2 = q0.w;
    const float r1r1 = r1 * r1;
    const float r2r2 = r2 * r2;
    const float tth = (r2 - r1) / (z2 - z1);
    const float tth2 = tth * tth;
    const float z0 = (z2 * r1 - z1 * r2) / (r1 - r2);
    const float idz = 1.f / d.z;
    float t_cap1 = d.z == 0.f ? OA_RT_DEFAULT_MAX : (z1 - o.z) * idz;
    float t_cap2 = d.z == 0.f ? OA_RT_DEFAULT_MAX : (z2 - o.z) * idz;
    const float rr_cap1 = (o.x + t_cap1 * d.x) * (o.x + t_cap1 * d.x) + (o.y + t_cap1 * d.y) * (o.y + t_cap1 * d.y);
    const float rr_cap2 = (o.x + t_cap2 * d.x) * (o.x + t_cap2 * d.x) + (o.y + t_cap2 * d.y) * (o.y + t_cap2 * d.y);
    t_cap1 = rr_cap1 < r1r1 && t_cap1 > t_min ? t_cap1 : OA_RT_DEFAULT_MAX;
    t_cap2 = rr_cap2 < r2r2 && t_cap2 > t_min ? t_cap2 : OA_RT_DEFAULT_MAX;
    const float c2 = d.x * d.x + d.y * d.y - d.z * d.z * tth2;
    const float c1 = o.x * d.x + o.y * d.y - (o.z - z0) * d.z * tth2;
    const float c0 = o.x * o.x + o.y * o.y - (o.z - z0) * (o.z - z0) * tth2;
    float t_near, t_far, sdisc;
    rq_roots_disq(OA_RT_DEFAULT_MAX, t_near, t_far, sdisc, c2, c1, c0);
    const float z_near = o.z + t_near * d.z;
    const float z_far = o.z + t_far * d.z;
    t_near = z_near > z1 && z_near < z2 && t_near > t_min ? t_near : OA_RT_DEFAULT_MAX;
    t_far = z_far > z1 && z_far < z2 && t_far > t_min ? t_far : OA_RT_DEFAULT_MAX;
    const float t_cand = fmin4(t_near, t_far, t_cap1, t_cap2);
    const bool valid = t_cand > t_min && t_cand < OA_RT_DEFAULT_MAX;
    if (valid) {
        if (t_cand == t_cap1 || t_cand == t_cap2) {
            is.x = 0.f; is.y = 0.f; is.z = t_cand == t_cap2 ? 1.f : -1.f;
        } else {
            const v3 n = normalize(mk(o.x + t_cand * d.x, o.y + t_cand * d.y, (z0 - (o.z + t_cand * d.z)) * tth2));
            is.x = n.x; is.y = n.y; is.z = n.z;
        }
        is.w = t_cand;
    }
    return valid;
}

/* intersection of plane half-spaces plan[planeIdx + i], i < planeNum
   (CSG/csg_intersect_leaf_convexpolyhedron.h:48-106) */
OA_D bool leaf_convexpolyhedron(Isect& is, float4 q0, const float4* plan, float t_min, v3 o, v3 d)
{
    const uint32_t planeIdx = __float_as_uint(q0.x), planeNum = __float_as_uint(q0.y);
    float t0 = -__uint_as_float(0x7f800000u), t1 = __uint_as_float(0x7f800000u);
    v3 t0_normal = mk(0.f, 0.f, 0.f), t1_normal = mk(0.f, 0.f, 0.f);
    for (uint32_t i = 0; i < planeNum; ++i) {
        const float4 pl = plan[planeIdx + i];
        const v3 n = mk(pl.x, pl.y, pl.z);
        const float dplane = pl.w;
        const float nd = dot(n, d);
        const float no = dot(n, o);
        const float dist = no - dplane;
        const float t_cand = -dist / nd;
        const bool parallel_inside = nd == 0.f && dist < 0.f;
        const bool parallel_outside = nd == 0.f && dist > 0.f;
        if (parallel_inside) continue;
        if (parallel_outside) return false;
        if (nd < 0.f) {
            if (t_cand > t0) { t0 = t_cand; t0_normal = n; }
        } else {
            if (t_cand < t1) { t1 = t_cand; t1_normal = n; }
        }
    }
    const bool valid = t0 < t1;
    if (valid) {
        if (t0 > t_min) {
            is.x = t0_normal.x; is.y = t0_normal.y; is.z = t0_normal.z; is.w = t0;
        } else if (t1 > t_min) {
            is.x = t1_normal.x; is.y = t1_normal.y; is.z = t1_normal.z; is.w = t1;
        }
    }
    return valid;
}

/* one-sheet hyperboloid with caps (CSG/csg_intersect_leaf_hyperboloid.h:34-106) */
OA_D bool leaf_hyperboloid(Isect& is, float4 q0, float t_min, v3 o, v3 s)
{
    const float zero = 0.f, one = 1.f;
    const float r0 = q0.x, zf = q0.y, z1 = q0.z, z2 = q0.w;
    const float rr0 = r0 * r0;
    const float z1s = z1 / zf;
    const float z2s = z2 / zf;
    const float rr1 = rr0 * (z1s * z1s + one);
    const float rr2 = rr0 * (z2s * z2s + one);
    const float A = -rr0 / (zf * zf);
    const float B = -rr0;
    const float dd = s.x * s.x + s.y * s.y + A * s.z * s.z;
    const float b = o.x * s.x + o.y * s.y + A * o.z * s.z;
    const float c = o.x * o.x + o.y * o.y + A * o.z * o.z + B;
    float t1hyp, t2hyp, sdisc;
    const float disc = b * b - dd * c;
    rq_roots(t1hyp, t2hyp, sdisc, dd, b, c);
    const float h1z = o.z + t1hyp * s.z;
    const float h2z = o.z + t2hyp * s.z;
    const float osz = one / s.z;
    const float t2cap = (z2 - o.z) * osz;
    const float t1cap = (z1 - o.z) * osz;
    const v3 c1 = o + t1cap * s;
    const v3 c2 = o + t2cap * s;
    const float crr1 = c1.x * c1.x + c1.y * c1.y;
    const float crr2 = c2.x * c2.x + c2.y * c2.y;
    const float ca = t1hyp > t_min && disc > zero && h1z > z1 && h1z < z2 ? t1hyp : OA_RT_DEFAULT_MAX;
    const float cb = t2hyp > t_min && disc > zero && h2z > z1 && h2z < z2 ? t2hyp : OA_RT_DEFAULT_MAX;
    const float cc = t2cap > t_min && crr2 < rr2 ? t2cap : OA_RT_DEFAULT_MAX;
    const float cd = t1cap > t_min && crr1 < rr1 ? t1cap : OA_RT_DEFAULT_MAX;
    const float t_cand = fmin4(ca, cb, cc, cd);
    const bool valid = t_cand > t_min && t_cand < OA_RT_DEFAULT_MAX;
    if (valid) {
        is.w = t_cand;
        if (t_cand == t1hyp || t_cand == t2hyp) {
            const v3 p = o + t_cand * s;
            const v3 n = normalize(mk(p.x, p.y, A * p.z));
            is.x = n.x; is.y = n.y; is.z = n.z;
        } else {
            is.x = zero; is.y = zero; is.z = t_cand == t1cap ? -one : one;
        }
    }
    return valid;
}

/* wedge between the half-planes at phi0, phi1 (intersect_leaf_phicut_simple,
   CSG/csg_intersect_leaf_phicut.h); a miss that exits at infinity sets y = -0 */
OA_D bool leaf_phicut(Isect& is, float4 q0, float t_min, v3 o, v3 d)
{
    const float cosPhi0 = q0.x, sinPhi0 = q0.y, cosPhi1 = q0.z, sinPhi1 = q0.w;
    const float d_n0 = d.x * sinPhi0 + d.y * (-cosPhi0);
    const float d_n1 = d.x * (-sinPhi1) + d.y * (cosPhi1);
    const float o_n0 = o.x * sinPhi0 + o.y * (-cosPhi0);
    const float o_n1 = o.x * (-sinPhi1) + o.y * (cosPhi1);
    float t0 = d_n0 == 0.f ? t_min : -o_n0 / d_n0;
    float t1 = d_n1 == 0.f ? t_min : -o_n1 / d_n1;
    const float PR = d_n0 == 0.f ? -o_n0 : -d_n0;
    const float QR = d_n1 == 0.f ? -o_n1 : -d_n1;
    const float PQ = cosPhi0 * sinPhi1 - cosPhi1 * sinPhi0;
    const bool unbounded_exit = PQ >= 0.f ? (PR >= 0.f && QR <= 0.f) : (PR >= 0.f || QR <= 0.f);
    const float side0 = o.x * cosPhi0 + o.y * sinPhi0 + (d.x * cosPhi0 + d.y * sinPhi0) * t0;
    const float side1 = o.x * cosPhi1 + o.y * sinPhi1 + (d.x * cosPhi1 + d.y * sinPhi1) * t1;
    if (side0 < 0.f) t0 = t_min;
    if (side1 < 0.f) t1 = t_min;
    const float t_near = fminf(t0, t1);
    const float t_far = fmaxf(t0, t1);
    const float t_cand = t_near > t_min ? t_near : (t_far > t_min ? t_far : t_min);
    const bool valid = t_cand > t_min;
    if (valid) {
        is.x = t_cand == t1 ? -sinPhi1 : sinPhi0;
        is.y = t_cand == t1 ? cosPhi1 : -cosPhi0;
        is.z = 0.f;
        is.w = t_cand;
    } else if (unbounded_exit) {
        is.y = -is.y;
    }
    return valid;
}

/* the region n.x <= w (CSG/csg_intersect_leaf_halfspace.h) */
OA_D bool leaf_halfspace(Isect& is, float4 q0, float t_min, v3 o, v3 d)
{
    const v3 n = mk(q0.x, q0.y, q0.z);
    const float w = q0.w;
    const float on = dot(o, n);
    const float dn = dot(d, n);
    const float on_w = on - w;
    const float adn = fabsf(dn);
    const float eps = 1e-9f;
    const bool inside = on_w < -eps;
    const float t = adn > 0.f ? -on_w / dn : t_min;
    const bool valid = t > t_min;
    if (valid) {
        is.x = n.x; is.y = n.y; is.z = n.z; is.w = t;
    } else if (inside) {
        is.y = -0.f;
    }
    return valid;
}

/* intersect_leaf (CSG/csg_intersect_leaf.h:174-324). Returns false for a miss.
   EXT: also the cone, convexpolyhedron, hyperboloid, phicut and halfspace leaves
   (a separate kernel instantiation, chosen when the geometry holds them). */
/* a leaf node given by its parameter quads q0, q1 and its q3.z (typecode), q3.w (transform | complement) */
template <bool EXT>
OA_D bool intersect_leaf_q(Isect& is, float4 q0, float4 q1, uint32_t typecode, uint32_t gtw, const float4* itra,
                           const float4* plan, float t_min, v3 ro, v3 rd)
{
    is.x = 0.f; is.y = 0.f; is.z = 0.f; is.w = 0.f;
    const uint32_t gt = gtw & 0x7fffffffu;
    const bool complement = (gtw & 0x80000000u) != 0u;
    const float4* q = gt > 0u ? itra + 4u * (gt - 1u) : nullptr;
    v3 origin = q ? q_right(q, ro, 1.f) : ro;
    v3 direction = q ? q_right(q, rd, 0.f) : rd;
    bool valid = false;
    switch (typecode) {
        case TC_SPHERE:   valid = leaf_sphere(is, q0, t_min, origin, direction); break;
        case TC_ZSPHERE:  valid = leaf_zsphere(is, q0, q1, t_min, origin, direction); break;
        case TC_CYLINDER: valid = leaf_cylinder(is, q0, q1, t_min, origin, direction); break;
        case TC_BOX3:     valid = leaf_box3(is, q0, t_min, origin, direction); break;
        default:
            if constexpr (EXT) {
                switch (typecode) {
                    case TC_CONE:     valid = leaf_cone(is, q0, t_min, origin, direction); break;
                    case TC_CONVEXPOLYHEDRON: valid = leaf_convexpolyhedron(is, q0, plan, t_min, origin, direction); break;
                    case TC_HYPERBOLOID: valid = leaf_hyperboloid(is, q0, t_min, origin, direction); break;
                    case TC_PHICUT:   valid = leaf_phicut(is, q0, t_min, origin, direction); break;
                    case TC_HALFSPACE: valid = leaf_halfspace(is, q0, t_min, origin, direction); break;
                    default: break;
                }
            }
            break;
    }
    if (valid && q) q_left_inplace(q, is, 0.f);
    if (complement) {
        is.x = valid ? -is.x : -0.f;
        is.y = valid ? -is.y : is.y;
        is.z = valid ? -is.z : is.z;
    }
    return valid;
}

template <bool EXT>
OA_D bool intersect_leaf(Isect& is, const float4* nd, const float4* itra, const float4* plan, float t_min, v3 ro, v3 rd)
{
    const float4 q3 = nd[3];
    return intersect_leaf_q<EXT>(is, nd[0], nd[1], __float_as_uint(q3.z), __float_as_uint(q3.w), itra, plan, t_min,
                                 ro, rd);
}

/* ------------------------------------------------------------- CSG tree */

/* Kensler decision table packed as the reference LUT words (CSG/csg_classify.h):
   nibble (3*stateA+stateB) of word [op-1], for A closer / B closer. */
__constant__ static const uint32_t kLutA[3] = { 0x22121141u, 0x00014014u, 0x00141141u };
__constant__ static const uint32_t kLutB[3] = { 0x22115122u, 0x00022055u, 0x00133155u };

enum : int { CT_MISS = 0, CT_RETURN_A = 1, CT_RETURN_B = 2, CT_RETURN_FLIP_B = 3, CT_LOOP_A = 4, CT_LOOP_B = 5 };
enum : int { ST_ENTER = 0, ST_EXIT = 1, ST_MISS = 2 };

OA_D int classify(const Isect& e, v3 dir, float tmin)
{
    return fabsf(e.w) > tmin ? (e.x * dir.x + e.y * dir.y + e.z * dir.z < 0.f ? ST_ENTER : ST_EXIT) : ST_MISS;
}

/* one operator node's decision from its left and right child results: a
   return (miss, A, B, flipped B) or a loop (advance A or B and re-evaluate) */
OA_D int csg_decide(uint32_t typecode, const Isect& L, const Isect& R, v3 rd, float tmin)
{
    int l_state = classify(L, rd, tmin);
    int r_state = classify(R, rd, tmin);
    const float t_left = fabsf(L.w);
    const float t_right = fabsf(R.w);
    bool leftIsCloser = t_left <= t_right;
    const bool l_promote = l_state == ST_MISS && (sgnbit(L.x) || sgnbit(L.y));
    const bool r_promote = r_state == ST_MISS && (sgnbit(R.x) || sgnbit(R.y));
    if (r_promote) { r_state = ST_EXIT; leftIsCloser = true; }
    if (l_promote) { l_state = ST_EXIT; leftIsCloser = false; }
    const uint32_t off = 3u * (uint32_t)l_state + (uint32_t)r_state;
    const uint32_t word = (typecode >= TC_UNION && typecode <= TC_DIFFERENCE)
                        ? (leftIsCloser ? kLutA[typecode - 1u] : kLutB[typecode - 1u]) : 0u;
    return off < 8u ? (int)((word >> (4u * off)) & 0xfu) : CT_MISS;
}

/* the intersect a return decision passes up (w keeps its magnitude) */
OA_D Isect csg_result(int ctrl, const Isect& L, const Isect& R)
{
    Isect res;
    if (ctrl == CT_MISS) { res.x = 0.f; res.y = 0.f; res.z = 0.f; res.w = 0.f; }
    else res = pick(ctrl == CT_RETURN_A, L, R);
    if (ctrl == CT_RETURN_FLIP_B) { res.x = -res.x; res.y = -res.y; res.z = -res.z; }
    return res;
}

OA_D uint32_t tdepth(uint32_t i) { return 31u - (uint32_t)__builtin_clz(i); }
OA_D uint32_t postorder_next(uint32_t i, uint32_t elev) { return (i & 1u) ? i >> 1 : (i << elev) + (1u << elev); }

/* ------------------------------------------------------ CSG list nodes */

/* List nodes keep subNum (q0.u.x) leaves at root + subOffset (q0.u.y) + i,
   counted from the prim's root node (CSG/CSGNode.h:187-211). They occur only in
   the EXT instantiation and run out of line: the contiguous case keeps small
   per-lane arrays of enter distances. An invalid result leaves the intersect
   zero, as the reference's list functions leave it untouched. */
#define OA_LIST_EPSILON 0.0001f
#define OA_LIST_MAX_ENTER 8

struct ListHit { float x, y, z, w; int valid; };

OA_D ListHit list_hit(const Isect& e, bool valid)
{
    ListHit h;
    h.x = valid ? e.x : 0.f; h.y = valid ? e.y : 0.f; h.z = valid ? e.z : 0.f; h.w = valid ? e.w : 0.f;
    h.valid = valid ? 1 : 0;
    return h;
}

/* intersect_node_contiguous (CSG/csg_intersect_node.h:392-650) */
__device__ static __attribute__((noinline)) ListHit list_contiguous(const float4* nd, const float4* root,
                                                                    const float4* itra, const float4* plan,
                                                                    float t_min, v3 ro, v3 rd)
{
    const int num_sub = (int)__float_as_uint(nd[0].x);
    const int off = (int)__float_as_uint(nd[0].y);
    Isect nearest_enter; nearest_enter.x = 0.f; nearest_enter.y = 0.f; nearest_enter.z = 0.f; nearest_enter.w = OA_RT_DEFAULT_MAX;
    Isect farthest_exit; farthest_exit.x = 0.f; farthest_exit.y = 0.f; farthest_exit.z = 0.f; farthest_exit.w = t_min;
    Isect sub;
    int exit_count = 0;
    for (int i = 0; i < num_sub; ++i) {
        if (!intersect_leaf<true>(sub, root + 4 * (off + i), itra, plan, t_min, ro, rd)) continue;
        const int st = classify(sub, rd, t_min);
        if (st == ST_ENTER) { if (sub.w < nearest_enter.w) nearest_enter = sub; }
        else if (st == ST_EXIT) exit_count += 1;
    }
    if (exit_count == 0)
        return list_hit(nearest_enter, nearest_enter.w > t_min && nearest_enter.w < OA_RT_DEFAULT_MAX);
    int enter_count = 0;
    float enter[OA_LIST_MAX_ENTER];
    int aux[OA_LIST_MAX_ENTER], idx[OA_LIST_MAX_ENTER];
    for (int isub = 0; isub < num_sub; ++isub) {
        if (!intersect_leaf<true>(sub, root + 4 * (off + isub), itra, plan, t_min, ro, rd)) continue;
        const int st = classify(sub, rd, t_min);
        if (st == ST_ENTER) {
            if (enter_count < OA_LIST_MAX_ENTER) {        /* num_sub <= 8 is checked at create */
                aux[enter_count] = isub; idx[enter_count] = enter_count; enter[enter_count] = sub.w;
                enter_count += 1;
            }
        } else if (st == ST_EXIT) {
            exit_count += 1;
            if (sub.w > farthest_exit.w) farthest_exit = sub;
        }
    }
    for (int i = 1; i < enter_count; ++i) {             /* indirect insertion sort by enter distance */
        const int key = idx[i];
        int j = i - 1;
        while (j >= 0 && enter[idx[j]] > enter[key]) { idx[j + 1] = idx[j]; j -= 1; }
        idx[j + 1] = key;
    }
    for (int i = 0; i < enter_count; ++i) {
        const float tminAdvanced = enter[idx[i]] + OA_LIST_EPSILON;
        const int isub = aux[idx[i]];
        if (tminAdvanced < farthest_exit.w) {
            if (intersect_leaf<true>(sub, root + 4 * (off + isub), itra, plan, tminAdvanced, ro, rd) &&
                classify(sub, rd, tminAdvanced) == ST_EXIT) {
                exit_count += 1;
                if (sub.w > farthest_exit.w) farthest_exit = sub;
            }
        }
    }
    return list_hit(farthest_exit, exit_count > 0 && farthest_exit.w > t_min);
}

/* intersect_node_discontiguous (CSG/csg_intersect_node.h:654-690) */
__device__ static __attribute__((noinline)) ListHit list_discontiguous(const float4* nd, const float4* root,
                                                                       const float4* itra, const float4* plan,
                                                                       float t_min, v3 ro, v3 rd)
{
    const int num_sub = (int)__float_as_uint(nd[0].x);
    const int off = (int)__float_as_uint(nd[0].y);
    Isect closest; closest.x = 0.f; closest.y = 0.f; closest.z = 0.f; closest.w = OA_RT_DEFAULT_MAX;
    Isect sub;
    for (int i = 0; i < num_sub; ++i)
        if (intersect_leaf<true>(sub, root + 4 * (off + i), itra, plan, t_min, ro, rd) && sub.w < closest.w)
            closest = sub;
    return list_hit(closest, closest.w < OA_RT_DEFAULT_MAX);
}

/* intersect_node_overlap (CSG/csg_intersect_node.h:819-905) */
__device__ static __attribute__((noinline)) ListHit list_overlap(const float4* nd, const float4* root,
                                                                 const float4* itra, const float4* plan,
                                                                 float t_min, v3 ro, v3 rd)
{
    const uint32_t num_sub = __float_as_uint(nd[0].x);
    const uint32_t off = __float_as_uint(nd[0].y);
    Isect farthest_enter; farthest_enter.x = 0.f; farthest_enter.y = 0.f; farthest_enter.z = 0.f; farthest_enter.w = t_min;
    Isect nearest_exit; nearest_exit.x = 0.f; nearest_exit.y = 0.f; nearest_exit.z = 0.f; nearest_exit.w = OA_RT_DEFAULT_MAX;
    Isect sub;
    uint32_t enter_count = 0u, exit_count = 0u;
    for (uint32_t isub = 0; isub < num_sub; ++isub) {
        const float4* sn = root + 4u * (off + isub);
        if (!intersect_leaf<true>(sub, sn, itra, plan, t_min, ro, rd)) continue;
        const int st = classify(sub, rd, t_min);
        if (st == ST_ENTER) {
            enter_count += 1u;
            if (sub.w > farthest_enter.w) farthest_enter = sub;
            const float tminAdvanced = sub.w + OA_LIST_EPSILON;
            if (intersect_leaf<true>(sub, sn, itra, plan, tminAdvanced, ro, rd) &&
                classify(sub, rd, tminAdvanced) == ST_EXIT) {
                exit_count += 1u;
                if (sub.w < nearest_exit.w) nearest_exit = sub;
            }
        } else if (st == ST_EXIT) {
            exit_count += 1u;
            if (sub.w < nearest_exit.w) nearest_exit = sub;
        }
    }
    const bool overlap_all = farthest_enter.w < nearest_exit.w && max(enter_count, exit_count) == num_sub;
    if (overlap_all && farthest_enter.w > t_min && farthest_enter.w < OA_RT_DEFAULT_MAX) return list_hit(farthest_enter, true);
    if (overlap_all && nearest_exit.w > t_min && nearest_exit.w < OA_RT_DEFAULT_MAX) return list_hit(nearest_exit, true);
    return list_hit(sub, false);
}

/* intersect_node (CSG/csg_intersect_node.h:913-937): a list node (EXT only) or a leaf */
template <bool EXT>
OA_D bool intersect_node(Isect& is, const float4* nd, const float4* root, const float4* itra, const float4* plan,
                         float t_min, v3 ro, v3 rd)
{
    if constexpr (EXT) {
        const uint32_t typecode = __float_as_uint(nd[3].z);
        if (typecode >= TC_CONTIGUOUS && typecode <= TC_OVERLAP) {
            const ListHit h = typecode == TC_CONTIGUOUS ? list_contiguous(nd, root, itra, plan, t_min, ro, rd)
                            : typecode == TC_DISCONTIGUOUS ? list_discontiguous(nd, root, itra, plan, t_min, ro, rd)
                            : list_overlap(nd, root, itra, plan, t_min, ro, rd);
            is.x = h.x; is.y = h.y; is.z = h.z; is.w = h.w;
            return h.valid != 0;
        }
    }
    return intersect_leaf<EXT>(is, nd, itra, plan, t_min, ro, rd);
}

/*
 * Device stacks. The evaluation of a complete tree of height h never holds
 * more than h+1 intersects (one pending left result per ancestor, plus the
 * loop's "other" and the re-evaluated subtree), so trees of height <= 3 use a
 * 4-deep stack and deeper ones (<= OA_CSG_MAX_HEIGHT, checked at create) an
 * 8-deep one. Both stacks are shift registers with the top at index 0: every
 * index is a compile-time constant, so they live in VGPRs, not scratch.
 * The tranche stack keeps the reference's 4 entries and overflow behaviour
 * (CSG/csg_tranche.h TRANCHE_STACK_SIZE).
 */
#define OA_TRANCHE_STACK 4

/* intersect_tree (CSG/csg_intersect_tree.h:277-672) */
template <int D, bool EXT>
OA_D bool intersect_tree_t(Isect& out, const float4* root, const float4* itra, const float4* plan, float t_min,
                           v3 ro, v3 rd)
{
    const uint32_t numNode = __float_as_uint(root[0].x);
    const uint32_t height = (uint32_t)(__builtin_ffs((int)numNode + 1) - 2);
    const float propagate_epsilon = 0.0001f;
    bool ierr = false;
    uint32_t trs[OA_TRANCHE_STACK];
    float trt[OA_TRANCHE_STACK];
    int trn = 1;
    trs[0] = (1u << height);          /* slice: begin | end<<8, fullTree end = 0 */
    trt[0] = t_min;
#pragma unroll
    for (int i = 1; i < OA_TRANCHE_STACK; ++i) { trs[i] = 0u; trt[i] = 0.f; }
    Isect st[D];
#pragma unroll
    for (int i = 0; i < D; ++i) { st[i].x = 0.f; st[i].y = 0.f; st[i].z = 0.f; st[i].w = 0.f; }
    int sn = 0;

    while (trn > 0) {
        const uint32_t slice = trs[0];
        const float tmin = trt[0];
#pragma unroll
        for (int i = 0; i < OA_TRANCHE_STACK - 1; ++i) { trs[i] = trs[i + 1]; trt[i] = trt[i + 1]; }
        trn--;
        const uint32_t beginIdx = slice & 0xffu;
        const uint32_t endIdx = (slice >> 8) & 0xffu;
        uint32_t nodeIdx = beginIdx;
        while (nodeIdx != endIdx) {
            const uint32_t depth = tdepth(nodeIdx);
            const uint32_t elevation = height - depth;
            const float4* nd = root + 4u * (nodeIdx - 1u);
            const uint32_t typecode = __float_as_uint(nd[3].z);
            if (typecode == TC_ZERO) { nodeIdx = postorder_next(nodeIdx, elevation); continue; }
            if (typecode >= TC_NODE) {
                Isect e;
                OA_COUNT(OA_ST_LEAF);
                OA_CYC_MARK(c_tleaf);
                intersect_node<EXT>(e, nd, root, itra, plan, tmin, ro, rd);
                OA_CYC_ADD(OA_ST_C_TLEAF, c_tleaf);
                e.w = copysignf(e.w, nodeIdx % 2u == 0u ? -1.f : 1.f);
                if (sn >= D) { ierr = true; break; }
#pragma unroll
                for (int i = D - 1; i > 0; --i) st[i] = st[i - 1];
                st[0] = e;
                sn++;
            } else {
                if (sn < 2) { ierr = true; break; }
                const bool firstLeft = sgnbit(st[0].w);
                const bool secondLeft = sgnbit(st[1].w);
                if (!(firstLeft ^ secondLeft)) { ierr = true; break; }
                const Isect L = pick(firstLeft, st[0], st[1]);
                const Isect R = pick(firstLeft, st[1], st[0]);
                const int ctrl = csg_decide(typecode, L, R, rd, tmin);
                /* pop two, push one: st[0] = result, entries below move up by one */
#pragma unroll
                for (int i = 1; i < D - 1; ++i) st[i] = st[i + 1];
                sn--;
                if (ctrl < CT_LOOP_A) {
                    Isect res = csg_result(ctrl, L, R);
                    res.w = copysignf(res.w, nodeIdx % 2u == 0u ? -1.f : 1.f);
                    st[0] = res;
                } else {
                    const bool loopA = ctrl == CT_LOOP_A;
                    const uint32_t leftIdx = 2u * nodeIdx;
                    const uint32_t rightIdx = leftIdx + 1u;
                    const float tminAdvanced = fabsf(loopA ? L.w : R.w) + propagate_epsilon;
                    st[0] = pick(loopA, R, L);
                    const uint32_t endTree = nodeIdx | (endIdx << 8);
                    const uint32_t leftTree = (leftIdx << (elevation - 1u)) | ((rightIdx << (elevation - 1u)) << 8);
                    const uint32_t rightTree = (rightIdx << (elevation - 1u)) | (nodeIdx << 8);
                    if (trn > OA_TRANCHE_STACK - 2) { ierr = true; break; }
#pragma unroll
                    for (int i = OA_TRANCHE_STACK - 1; i > 1; --i) { trs[i] = trs[i - 2]; trt[i] = trt[i - 2]; }
                    trs[1] = endTree;  trt[1] = tmin;
                    trs[0] = loopA ? leftTree : rightTree;  trt[0] = tminAdvanced;
                    trn += 2;
                    break;
                }
            }
            nodeIdx = postorder_next(nodeIdx, elevation);
        }
        if (ierr) break;
    }
    out.x = 0.f; out.y = 0.f; out.z = 0.f; out.w = 0.f;
    if (sn == 1) out = st[0];
    return out.w > 0.f;
}

/*
 * Trees of 3 or 7 nodes whose evaluation never loops. Each child of the root is
 * a leaf or an operator over two leaves, so the postorder algorithm above,
 * when no node decides to loop, evaluates every leaf once at t_min and
 * combines bottom-up in one tranche: this does the same in straight-line
 * code, with one leaf call site and no stacks. The sign that the general
 * algorithm puts on w to tell left from right is not needed here (every use
 * of w takes its magnitude) and the root's result carries |w| in both.
 * Returns -1 for any other shape or at the first loop decision: the caller
 * then runs the general algorithm from the start.
 */
template <bool EXT>
OA_D int intersect_tree_small(Isect& out, const float4* root, const float4* itra, const float4* plan, float t_min,
                              v3 ro, v3 rd)
{
    const uint32_t numNode = __float_as_uint(root[0].x);
    const uint32_t tc1 = __float_as_uint(root[3].z);
    if ((numNode != 3u && numNode != 7u) || tc1 < TC_UNION || tc1 > TC_DIFFERENCE) return -1;
    Isect sub[2];
    sub[0].x = sub[0].y = sub[0].z = sub[0].w = 0.f;
    sub[1] = sub[0];
    bool general = false;
#pragma unroll 1
    for (uint32_t s = 2u; s <= 3u; ++s) {
        const uint32_t tcs = __float_as_uint(root[4u * (s - 1u) + 3u].z);
        const bool sub_leaf = tcs >= TC_NODE;
        if (sub_leaf) {
            /* a leaf child of a 7-node root stands over two zero nodes */
            if (numNode == 7u && (__float_as_uint(root[4u * (2u * s - 1u) + 3u].z) != TC_ZERO ||
                                  __float_as_uint(root[4u * (2u * s) + 3u].z) != TC_ZERO)) { general = true; break; }
        } else if (numNode == 3u || tcs < TC_UNION || tcs > TC_DIFFERENCE) { general = true; break; }
        Isect x[2];
        x[0].x = x[0].y = x[0].z = x[0].w = 0.f;
        x[1] = x[0];
#pragma unroll 1
        for (uint32_t j = 0u; j < (sub_leaf ? 1u : 2u); ++j) {
            const uint32_t k = sub_leaf ? s : 2u * s + j;
            const float4* nd = root + 4u * (k - 1u);
            if (__float_as_uint(nd[3].z) < TC_NODE) { general = true; break; }
            Isect e;
            OA_COUNT(OA_ST_LEAF);
            intersect_node<EXT>(e, nd, root, itra, plan, t_min, ro, rd);
            x[1] = pick(j == 0u, x[1], e);
            x[0] = pick(j == 0u, e, x[0]);
        }
        if (general) break;
        Isect r = x[0];
        if (!sub_leaf) {
            const int ctrl = csg_decide(tcs, x[0], x[1], rd, t_min);
            if (ctrl >= CT_LOOP_A) { general = true; break; }
            r = csg_result(ctrl, x[0], x[1]);
        }
        sub[1] = pick(s == 2u, sub[1], r);
        sub[0] = pick(s == 2u, r, sub[0]);
    }
    if (general) return -1;
    const int ctrl = csg_decide(tc1, sub[0], sub[1], rd, t_min);
    if (ctrl >= CT_LOOP_A) return -1;
    out = csg_result(ctrl, sub[0], sub[1]);
    out.w = fabsf(out.w);
    return out.w > 0.f ? 1 : 0;
}

/* the general algorithm out of line (stack depth D: 4 up to height 3, 8 deeper), result
   returned by value: it runs only for trees the straight-line evaluator hands over, so
   its code stays out of the propagation loop */
template <int D, bool EXT>
__device__ static __attribute__((noinline)) float4 intersect_tree_deep(const float4* root, const float4* itra,
                                                                      const float4* plan,
                                                                      float t_min, float ox, float oy, float oz,
                                                                      float dx, float dy, float dz)
{
    Isect r;
    intersect_tree_t<D, EXT>(r, root, itra, plan, t_min, mk(ox, oy, oz), mk(dx, dy, dz));
    return make_float4(r.x, r.y, r.z, r.w);
}

template <bool EXT>
OA_D bool intersect_prim(Isect& is, const float4* root, const float4* itra, const float4* plan, float t_min, v3 ro,
                         v3 rd)
{
    /* the root's quads loaded together: a leaf root needs no second round trip */
    const float4 q0 = root[0], q1 = root[1], q3 = root[3];
    const uint32_t typecode = __float_as_uint(q3.z);
    if (typecode >= TC_LEAF)
        return intersect_leaf_q<EXT>(is, q0, q1, typecode, __float_as_uint(q3.w), itra, plan, t_min, ro, rd);
    is.x = 0.f; is.y = 0.f; is.z = 0.f; is.w = 0.f;
    if (typecode < TC_NODE) {
        OA_COUNT(OA_ST_TREE);
        OA_WAVE_COUNT(OA_ST_W_TREE, OA_ST_L_TREE);
        OA_CYC_MARK(c_tree);
        const uint32_t numNode = __float_as_uint(root[0].x);
        int h = numNode <= 7u ? intersect_tree_small<EXT>(is, root, itra, plan, t_min, ro, rd) : -1;
        if (h < 0) {
            OA_COUNT(OA_ST_TREE_GENERAL);
            const float4 r = numNode <= 15u
                ? intersect_tree_deep<4, EXT>(root, itra, plan, t_min, ro.x, ro.y, ro.z, rd.x, rd.y, rd.z)
                : intersect_tree_deep<8, EXT>(root, itra, plan, t_min, ro.x, ro.y, ro.z, rd.x, rd.y, rd.z);
            is.x = r.x; is.y = r.y; is.z = r.z; is.w = r.w;
            h = is.w > 0.f ? 1 : 0;
        }
        OA_CYC_ADD(OA_ST_C_TREE, c_tree);
        return h != 0;
    }
    if constexpr (EXT) return intersect_node<true>(is, root, root, itra, plan, t_min, ro, rd);
    return false;
}

} // namespace oa
