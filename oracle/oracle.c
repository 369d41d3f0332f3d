/*
 * oracle.c : CPU restatement of the reference's optical-photon hot path.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h). Each function cites the reference
 * lines it follows. The arithmetic is written out in the reference's operand
 * order (a*b + c*d ... left to right, no contraction: built with
 * -ffp-contract=off) so that the HIP kernels, written independently with the
 * same operand order, agree bit for bit on the same random streams.
 */
#include <stdlib.h>
#include <string.h>
#include <math.h>
#include <stdint.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "oracle.h"
#include "../opticks_amd/csrc/oa_portable.h"

/* ------------------------------------------------------------- vectors */
/* Operand order follows sysrap/scuda.h (dot, cross, length, normalize). */

typedef struct { float x, y, z; } f3;
typedef struct { float x, y, z, w; } f4;

static inline f3 mk3(float x, float y, float z) { f3 r; r.x = x; r.y = y; r.z = z; return r; }
static inline f4 mk4(float x, float y, float z, float w) { f4 r; r.x = x; r.y = y; r.z = z; r.w = w; return r; }
static inline f3 add3(f3 a, f3 b) { return mk3(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline f3 sub3(f3 a, f3 b) { return mk3(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline f3 smul3(float s, f3 a) { return mk3(s * a.x, s * a.y, s * a.z); }   /* s*v */
static inline f3 mul3s(f3 a, float s) { return mk3(a.x * s, a.y * s, a.z * s); }   /* v*s */
static inline f3 neg3(f3 a) { return mk3(-a.x, -a.y, -a.z); }
static inline float dot3(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline f3 cross3(f3 a, f3 b)
{
    return mk3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
static inline float len3(f3 v) { return sqrtf(dot3(v, v)); }
static inline f3 norm3(f3 v) { float inv = 1.0f / sqrtf(dot3(v, v)); return mul3s(v, inv); }

#define PI_F 3.14159274101257324f   /* M_PIf, float(pi) */

/* -------------------------------------------------------------- flags */
/* sysrap/OpticksPhoton.h:22-44 */
enum {
    F_CERENKOV = 0x1 << 0, F_SCINTILLATION = 0x1 << 1, F_TORCH = 0x1 << 2,
    F_BULK_ABSORB = 0x1 << 3, F_BULK_REEMIT = 0x1 << 4, F_BULK_SCATTER = 0x1 << 5,
    F_SURFACE_DETECT = 0x1 << 6, F_SURFACE_ABSORB = 0x1 << 7,
    F_SURFACE_DREFLECT = 0x1 << 8, F_SURFACE_SREFLECT = 0x1 << 9,
    F_BOUNDARY_REFLECT = 0x1 << 10, F_BOUNDARY_TRANSMIT = 0x1 << 11
};
/* sysrap/sflow.h */
enum { C_UNDEFINED, C_BREAK, C_CONTINUE, C_BOUNDARY };
/* sysrap/smatsur.h:8-16 */
enum { EMS_MATERIAL = 0, EMS_NOSURFACE = 1, EMS_SURFACE = 2, EMS_ZPLUS_SENSOR_A = 3,
       EMS_ZPLUS_SENSOR_CUSTOMART = 4, EMS_ZMINUS = 5 };
/* sysrap/OpticksCSG.h */
enum { CSG_ZERO = 0, CSG_UNION = 1, CSG_INTERSECTION = 2, CSG_DIFFERENCE = 3, CSG_NODE = 11,
       CSG_CONTIGUOUS = 11, CSG_DISCONTIGUOUS = 12, CSG_OVERLAP = 13,
       CSG_LEAF = 101, CSG_SPHERE = 101, CSG_ZSPHERE = 103, CSG_CYLINDER = 105, CSG_CONE = 108, CSG_BOX3 = 110,
       CSG_CONVEXPOLYHEDRON = 112, CSG_HYPERBOLOID = 117, CSG_PHICUT = 121, CSG_HALFSPACE = 125 };
/* sysrap/OpticksGenstep.h:20-45 */
enum { GS_TORCH = 6, GS_SCINT_R4695 = 5, GS_CERENKOV = 15, GS_SCINTILLATION = 16, GS_G4CERENKOV_MODIFIED = 18 };
/* sysrap/storchtype.h:10-18; T_POINT_ISO is this project's isotropic point source */
enum { T_UNDEF, T_DISC, T_LINE, T_POINT, T_CIRCLE, T_RECTANGLE, T_SPHERE_MARSAGLIA, T_SPHERE,
       T_POINT_ISO = 100 };

/* ------------------------------------------------------------ photon */
/* sysrap/sphoton.h:171-246 */
typedef struct {
    f3 pos; float time;
    f3 mom; uint32_t hitcount_iindex;
    f3 pol; float wavelength;
    uint32_t orient_boundary_flag, identity, index, flagmask;
} oph;

static inline uint32_t ph_flag(const oph* p) { return p->orient_boundary_flag & 0xffffu; }
static inline uint32_t ph_boundary(const oph* p) { return (p->orient_boundary_flag & 0x7fff0000u) >> 16; }
static inline void ph_set_flag(oph* p, uint32_t flag)
{
    p->orient_boundary_flag = (p->orient_boundary_flag & 0xffff0000u) | (flag & 0xffffu);
    p->flagmask |= flag;
}
static inline void ph_zero_flags(oph* p)
{
    p->orient_boundary_flag = 0u; p->identity = 0u; p->index = 0u; p->flagmask = 0u; p->hitcount_iindex = 0u;
}
/* sphoton::set_prd (sysrap/sphoton.h:499-505) */
static inline void ph_set_prd(oph* p, uint32_t boundary, uint32_t identity, float orient, uint32_t iindex)
{
    p->orient_boundary_flag = (p->orient_boundary_flag & 0x8000ffffu) | ((boundary & 0x7fffu) << 16);
    p->identity = (p->identity & 0xff000000u) | (identity & 0x00ffffffu);
    p->orient_boundary_flag = (p->orient_boundary_flag & 0x7fffffffu) | ((orient < 0.f ? 0x1u : 0x0u) << 31);
    p->hitcount_iindex = 0x00010000u | (0x0000ffffu & iindex);
}
/* sphoton::set_index (sysrap/sphoton.h:230-237) */
static inline void ph_set_index(oph* p, uint64_t full)
{
    p->index = (uint32_t)(full & 0xffffffffu);
    p->identity = ((uint32_t)((full >> 32) & 0xffu) << 24) | (p->identity & 0xffffffu);
}

/* --------------------------------------------------------------- PRD */
/* sysrap/squad.h:180-250 quad2 */
typedef struct {
    f3 normal; float t;
    float lposcost, lposfphi;
    uint32_t iindex_identity, gp_boundary;
    uint32_t inst;          /* instance index of the hit (not in quad2) */
    float lx, ly;           /* local hit position x, y (lposfphi source); miss: 0, NaN */
} oprd;

/* ------------------------------------------------------------ random */

/* kind 0 Philox, 1 precooked uniforms (row of nk values per photon, the
   cursor wrapping modulo nk as s_seq::flat recycles, sysrap/s_seq.h:105-129),
   2 xorwow (oa_portable.h convention) */
typedef struct {
    int kind;
    oa_philox_stream ph;
    oa_xorwow_state xw;
    const float* pre;       /* precooked row for this photon */
    int64_t nk, cursor;
} orng;

/* the reference's rejection loops are unbounded; both sides stop after this
   many trials (unreachable for random streams and valid inputs) */
#define OA_MAX_TRIALS 65536

static inline float ou(orng* r)
{
    if (r->kind == 1) {
        float v = r->pre[r->cursor];
        r->cursor = r->cursor + 1 == r->nk ? 0 : r->cursor + 1;
        return v;
    }
    if (r->kind == 2) return oa_xorwow_uniform(&r->xw);
    return oa_rng_uniform(&r->ph);
}

/* ---------------------------------------------------------- geometry */


typedef struct { double lo[3], hi[3]; } dbox;

typedef struct {
    int left, right;        /* children, or -1 */
    int first, count;       /* leaf range into order[] */
    dbox box;
} bnode;

typedef struct {
    oracle_geom g;
    void* w2o;              /* n_inst w2o16 records */
    dbox* ibox;             /* n_inst world boxes (padded) */
    int32_t* inst_gas;      /* solid index per instance */
    uint32_t* inst_ident;   /* sensor_identifier+1 per instance */
    int32_t* solid_np;      /* numPrim per solid */
    int32_t* solid_po;      /* primOffset per solid */
    bnode* nodes; int n_nodes;
    int32_t* order;
} oworld;

static inline const uint32_t* nodeu(const float* node) { return (const uint32_t*)node; }

/* CSGNode field accessors (CSG/CSGNode.h:196-345) */
static inline uint32_t nd_typecode(const float* nd) { return nodeu(nd)[14]; }
static inline uint32_t nd_gtransformIdx(const float* nd) { return nodeu(nd)[15] & 0x7fffffffu; }
static inline int nd_complement(const float* nd) { return (nodeu(nd)[15] & 0x80000000u) != 0u; }
static inline uint32_t nd_boundary(const float* nd) { return nodeu(nd)[6]; }
static inline uint32_t nd_subNum(const float* nd) { return nodeu(nd)[0]; }
static inline uint32_t nd_subOffset(const float* nd) { return nodeu(nd)[1]; }

/* qat4::right_multiply (sysrap/sqat4.h:45-51): v^T M with w */
static inline f3 q_right_multiply(const float* q, f3 v, float w)
{
    f3 r;
    r.x = q[0] * v.x + q[4] * v.y + q[8] * v.z + q[12] * w;
    r.y = q[1] * v.x + q[5] * v.y + q[9] * v.z + q[13] * w;
    r.z = q[2] * v.x + q[6] * v.y + q[10] * v.z + q[14] * w;
    return r;
}
/* qat4::left_multiply_inplace (sysrap/sqat4.h:96-104) */
static inline void q_left_multiply_inplace(const float* q, f4* v, float w)
{
    float x = q[0] * v->x + q[1] * v->y + q[2] * v->z + q[3] * w;
    float y = q[4] * v->x + q[5] * v->y + q[6] * v->z + q[7] * w;
    float z = q[8] * v->x + q[9] * v->y + q[10] * v->z + q[11] * w;
    v->x = x; v->y = y; v->z = z;
}

/* ----------------------------------------------- robust quadratic roots */
/* CSG/csg_robust_quadratic_roots.h:104-136 */
static inline void robust_quadratic_roots(float* t1, float* t2, float* disc, float* sdisc,
                                          float d, float b, float c)
{
    *disc = b * b - d * c;
    *sdisc = *disc > 0.f ? sqrtf(*disc) : 0.f;
    float q = b > 0.f ? -(b + *sdisc) : -(b - *sdisc);
    float root1 = q / d;
    float root2 = c / q;
    *t1 = fminf(root1, root2);
    *t2 = fmaxf(root1, root2);
}
static inline void robust_quadratic_roots_disqualifying(float t_min, float* t1, float* t2, float* disc,
                                                        float* sdisc, float d, float b, float c)
{
    *disc = b * b - d * c;
    *sdisc = *disc > 0.f ? sqrtf(*disc) : 0.f;
    float q = b > 0.f ? -(b + *sdisc) : -(b - *sdisc);
    float root1 = *sdisc > 0.f ? q / d : t_min;
    float root2 = *sdisc > 0.f ? c / q : t_min;
    *t1 = fminf(root1, root2);
    *t2 = fmaxf(root1, root2);
}

/* ---------------------------------------------------------------- leaves */

/* CSG/csg_intersect_leaf_sphere.h:14-57 */
static void leaf_sphere(int* valid, f4* isect, const float* q0, float t_min, f3 o, f3 dir)
{
    f3 center = mk3(q0[0], q0[1], q0[2]);
    float radius = q0[3];
    f3 O = sub3(o, center);
    f3 D = dir;
    float b = dot3(O, D);
    float c = dot3(O, O) - radius * radius;
    float d = dot3(D, D);
    float root1, root2, disc, sdisc;
    robust_quadratic_roots(&root1, &root2, &disc, &sdisc, d, b, c);
    float t_cand = sdisc > 0.f ? (root1 > t_min ? root1 : root2) : t_min;
    *valid = t_cand > t_min;
    if (*valid) {
        isect->x = (O.x + t_cand * D.x) / radius;
        isect->y = (O.y + t_cand * D.y) / radius;
        isect->z = (O.z + t_cand * D.z) / radius;
        isect->w = t_cand;
    }
}

/* CSG/csg_intersect_leaf_zsphere.h:55-153 */
static void leaf_zsphere(int* valid, f4* isect, const float* q0, const float* q1, float t_min, f3 ro, f3 rd)
{
    f3 center = mk3(q0[0], q0[1], q0[2]);
    f3 O = sub3(ro, center);
    f3 D = rd;
    float radius = q0[3];
    float b = dot3(O, D);
    float c = dot3(O, O) - radius * radius;
    if (c > 0.f && b > 0.f) { *valid = 0; return; }
    float zmax = center.z + q1[1];
    float zmin = center.z + q1[0];
    float d = dot3(D, D);
    float t1sph, t2sph, disc, sdisc;
    robust_quadratic_roots(&t1sph, &t2sph, &disc, &sdisc, d, b, c);
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
    *valid = t_cand > t_min;
    if (*valid) {
        isect->w = t_cand;
        if (t_cand == t1sph || t_cand == t2sph) {
            isect->x = (O.x + t_cand * D.x) / radius;
            isect->y = (O.y + t_cand * D.y) / radius;
            isect->z = (O.z + t_cand * D.z) / radius;
        } else {
            isect->x = 0.f;
            isect->y = 0.f;
            isect->z = t_cand == t_PCAP ? -1.f : 1.f;
        }
    }
}

/* CSG/csg_intersect_leaf_cylinder.h:36-104 */
static void leaf_cylinder(int* valid, f4* isect, const float* q0, const float* q1, float t_min, f3 ro, f3 rd)
{
    float r = q0[3], z1 = q1[0], z2 = q1[1];
    float ox = ro.x, oy = ro.y, oz = ro.z, vx = rd.x, vy = rd.y, vz = rd.z;
    float r2 = r * r;
    float a = vx * vx + vy * vy;
    float b = ox * vx + oy * vy;
    float c = ox * ox + oy * oy - r2;
    float t_near, t_far, disc, sdisc;
    robust_quadratic_roots_disqualifying(t_min, &t_near, &t_far, &disc, &sdisc, a, b, c);
    float z_near = oz + t_near * vz;
    float z_far = oz + t_far * vz;
    float t_z1cap = (z1 - oz) / vz;
    float r2_z1cap = (ox + t_z1cap * vx) * (ox + t_z1cap * vx) + (oy + t_z1cap * vy) * (oy + t_z1cap * vy);
    float t_z2cap = (z2 - oz) / vz;
    float r2_z2cap = (ox + t_z2cap * vx) * (ox + t_z2cap * vx) + (oy + t_z2cap * vy) * (oy + t_z2cap * vy);
    float t_cand = INFINITY;
    if (t_near > t_min && z_near > z1 && z_near < z2 && t_near < t_cand) t_cand = t_near;
    if (t_far > t_min && z_far > z1 && z_far < z2 && t_far < t_cand) t_cand = t_far;
    if (t_z1cap > t_min && r2_z1cap <= r2 && t_z1cap < t_cand) t_cand = t_z1cap;
    if (t_z2cap > t_min && r2_z2cap <= r2 && t_z2cap < t_cand) t_cand = t_z2cap;
    *valid = t_cand > t_min && t_cand < INFINITY;
    if (*valid) {
        int sheet = (t_cand == t_near || t_cand == t_far);
        isect->x = sheet ? (ox + t_cand * vx) / r : 0.f;
        isect->y = sheet ? (oy + t_cand * vy) / r : 0.f;
        isect->z = sheet ? 0.f : (t_cand == t_z1cap ? -1.f : 1.f);
        isect->w = t_cand;
    }
}

/* CSG/csg_intersect_leaf_box3.h:60-157 (fullside dimensions, centred) */
static void leaf_box3(int* valid, f4* isect, const float* q0, float t_min, f3 ro, f3 rd)
{
    f3 bmin = mk3(-q0[0] / 2.f, -q0[1] / 2.f, -q0[2] / 2.f);
    f3 bmax = mk3(q0[0] / 2.f, q0[1] / 2.f, q0[2] / 2.f);
    f3 bcen = mk3(0.f, 0.f, 0.f);
    f3 idir = mk3(1.f / rd.x, 1.f / rd.y, 1.f / rd.z);
    f3 t0 = mk3((bmin.x - ro.x) * idir.x, (bmin.y - ro.y) * idir.y, (bmin.z - ro.z) * idir.z);
    f3 t1 = mk3((bmax.x - ro.x) * idir.x, (bmax.y - ro.y) * idir.y, (bmax.z - ro.z) * idir.z);
    f3 nr = mk3(fminf(t0.x, t1.x), fminf(t0.y, t1.y), fminf(t0.z, t1.z));
    f3 fr = mk3(fmaxf(t0.x, t1.x), fmaxf(t0.y, t1.y), fmaxf(t0.z, t1.z));
    float t_near = fmaxf(fmaxf(nr.x, nr.y), nr.z);
    float t_far = fminf(fminf(fr.x, fr.y), fr.z);
    int along_x = rd.x != 0.f && rd.y == 0.f && rd.z == 0.f;
    int along_y = rd.x == 0.f && rd.y != 0.f && rd.z == 0.f;
    int along_z = rd.x == 0.f && rd.y == 0.f && rd.z != 0.f;
    int in_x = ro.x > bmin.x && ro.x < bmax.x;
    int in_y = ro.y > bmin.y && ro.y < bmax.y;
    int in_z = ro.z > bmin.z && ro.z < bmax.z;
    int has_intersect;
    if (along_x) has_intersect = in_y && in_z;
    else if (along_y) has_intersect = in_x && in_z;
    else if (along_z) has_intersect = in_x && in_y;
    else has_intersect = (t_far > t_near && t_far > 0.f);
    *valid = 0;
    if (has_intersect) {
        float t_cand = t_min < t_near ? t_near : (t_min < t_far ? t_far : t_min);
        f3 p = sub3(add3(ro, smul3(t_cand, rd)), bcen);
        f3 pa = mk3(fabsf(p.x) / (bmax.x - bmin.x), fabsf(p.y) / (bmax.y - bmin.y), fabsf(p.z) / (bmax.z - bmin.z));
        f3 n = mk3(0.f, 0.f, 0.f);
        if (pa.x >= pa.y && pa.x >= pa.z) n.x = copysignf(1.f, p.x);
        else if (pa.y >= pa.x && pa.y >= pa.z) n.y = copysignf(1.f, p.y);
        else if (pa.z >= pa.x && pa.z >= pa.y) n.z = copysignf(1.f, p.z);
        if (t_cand > t_min) {
            *valid = 1;
            isect->x = n.x; isect->y = n.y; isect->z = n.z; isect->w = t_cand;
        }
    }
}

/* intersect_leaf (CSG/csg_intersect_leaf.h:174-324), non-HEISENBUG form */
#define RT_DEFAULT_MAX_F 1.e27f
static inline float fmin4(float a, float b, float c, float d) { return fminf(fminf(a, b), fminf(c, d)); }

/* csg_intersect_leaf_newcone.h:45-123 intersect_leaf_newcone */
static void leaf_cone(int* valid, f4* isect, const float* q0, float t_min, f3 o, f3 d)
{
    const float r1 = q0[0], z1 = q0[1], r2 = q0[2], z2 = q0[3];
    const float r1r1 = r1 * r1;
    const float r2r2 = r2 * r2;
    const float tth = (r2 - r1) / (z2 - z1);
    const float tth2 = tth * tth;
    const float z0 = (z2 * r1 - z1 * r2) / (r1 - r2);
    const float idz = 1.f / d.z;
    float t_cap1 = d.z == 0.f ? RT_DEFAULT_MAX_F : (z1 - o.z) * idz;
    float t_cap2 = d.z == 0.f ? RT_DEFAULT_MAX_F : (z2 - o.z) * idz;
    const float rr_cap1 = (o.x + t_cap1 * d.x) * (o.x + t_cap1 * d.x) + (o.y + t_cap1 * d.y) * (o.y + t_cap1 * d.y);
    const float rr_cap2 = (o.x + t_cap2 * d.x) * (o.x + t_cap2 * d.x) + (o.y + t_cap2 * d.y) * (o.y + t_cap2 * d.y);
    t_cap1 = rr_cap1 < r1r1 && t_cap1 > t_min ? t_cap1 : RT_DEFAULT_MAX_F;
    t_cap2 = rr_cap2 < r2r2 && t_cap2 > t_min ? t_cap2 : RT_DEFAULT_MAX_F;
    const float c2 = d.x * d.x + d.y * d.y - d.z * d.z * tth2;
    const float c1 = o.x * d.x + o.y * d.y - (o.z - z0) * d.z * tth2;
    const float c0 = o.x * o.x + o.y * o.y - (o.z - z0) * (o.z - z0) * tth2;
    float t_near, t_far, disc, sdisc;
    robust_quadratic_roots_disqualifying(RT_DEFAULT_MAX_F, &t_near, &t_far, &disc, &sdisc, c2, c1, c0);
    const float z_near = o.z + t_near * d.z;
    const float z_far = o.z + t_far * d.z;
    t_near = z_near > z1 && z_near < z2 && t_near > t_min ? t_near : RT_DEFAULT_MAX_F;
    t_far = z_far > z1 && z_far < z2 && t_far > t_min ? t_far : RT_DEFAULT_MAX_F;
    const float t_cand = fmin4(t_near, t_far, t_cap1, t_cap2);
    *valid = t_cand > t_min && t_cand < RT_DEFAULT_MAX_F;
    if (*valid) {
        if (t_cand == t_cap1 || t_cand == t_cap2) {
            isect->x = 0.f; isect->y = 0.f; isect->z = t_cand == t_cap2 ? 1.f : -1.f;
        } else {
            f3 n = norm3(mk3(o.x + t_cand * d.x, o.y + t_cand * d.y, (z0 - (o.z + t_cand * d.z)) * tth2));
            isect->x = n.x; isect->y = n.y; isect->z = n.z;
        }
        isect->w = t_cand;
    }
}

/* csg_intersect_leaf_convexpolyhedron.h:48-106; planes plan[planeIdx + i] (q0.u.x, q0.u.y) */
static void leaf_convexpolyhedron(int* valid, f4* isect, const float* nd, const float* plan, float t_min, f3 o, f3 d)
{
    const uint32_t planeIdx = nodeu(nd)[0], planeNum = nodeu(nd)[1];
    float t0 = -INFINITY, t1 = INFINITY;
    f3 t0_normal = mk3(0.f, 0.f, 0.f), t1_normal = mk3(0.f, 0.f, 0.f);
    for (uint32_t i = 0; i < planeNum; ++i) {
        const float* pl = plan + 4 * (planeIdx + i);
        f3 n = mk3(pl[0], pl[1], pl[2]);
        float dplane = pl[3];
        float ndd = dot3(n, d);
        float no = dot3(n, o);
        float dist = no - dplane;
        float t_cand = -dist / ndd;
        int parallel_inside = ndd == 0.f && dist < 0.f;
        int parallel_outside = ndd == 0.f && dist > 0.f;
        if (parallel_inside) continue;
        if (parallel_outside) { *valid = 0; return; }
        if (ndd < 0.f) {
            if (t_cand > t0) { t0 = t_cand; t0_normal = n; }
        } else {
            if (t_cand < t1) { t1 = t_cand; t1_normal = n; }
        }
    }
    *valid = t0 < t1;
    if (*valid) {
        if (t0 > t_min) {
            isect->x = t0_normal.x; isect->y = t0_normal.y; isect->z = t0_normal.z; isect->w = t0;
        } else if (t1 > t_min) {
            isect->x = t1_normal.x; isect->y = t1_normal.y; isect->z = t1_normal.z; isect->w = t1;
        }
    }
}

/* csg_intersect_leaf_hyperboloid.h:34-106 */
static void leaf_hyperboloid(int* valid, f4* isect, const float* q0, float t_min, f3 o, f3 s)
{
    const float zero = 0.f, one = 1.f;
    const float r0 = q0[0], zf = q0[1], z1 = q0[2], z2 = q0[3];
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
    float t1hyp, t2hyp, disc, sdisc;
    robust_quadratic_roots(&t1hyp, &t2hyp, &disc, &sdisc, dd, b, c);
    const float h1z = o.z + t1hyp * s.z;
    const float h2z = o.z + t2hyp * s.z;
    float osz = one / s.z;
    float t2cap = (z2 - o.z) * osz;
    float t1cap = (z1 - o.z) * osz;
    const f3 c1 = add3(o, smul3(t1cap, s));
    const f3 c2 = add3(o, smul3(t2cap, s));
    float crr1 = c1.x * c1.x + c1.y * c1.y;
    float crr2 = c2.x * c2.x + c2.y * c2.y;
    const float ca = t1hyp > t_min && disc > zero && h1z > z1 && h1z < z2 ? t1hyp : RT_DEFAULT_MAX_F;
    const float cb = t2hyp > t_min && disc > zero && h2z > z1 && h2z < z2 ? t2hyp : RT_DEFAULT_MAX_F;
    const float cc = t2cap > t_min && crr2 < rr2 ? t2cap : RT_DEFAULT_MAX_F;
    const float cd = t1cap > t_min && crr1 < rr1 ? t1cap : RT_DEFAULT_MAX_F;
    float t_cand = fmin4(ca, cb, cc, cd);
    *valid = t_cand > t_min && t_cand < RT_DEFAULT_MAX_F;
    if (*valid) {
        isect->w = t_cand;
        if (t_cand == t1hyp || t_cand == t2hyp) {
            const f3 p = add3(o, smul3(t_cand, s));
            f3 n = norm3(mk3(p.x, p.y, A * p.z));
            isect->x = n.x; isect->y = n.y; isect->z = n.z;
        } else {
            isect->x = zero; isect->y = zero; isect->z = t_cand == t1cap ? -one : one;
        }
    }
}

/* csg_intersect_leaf_phicut.h intersect_leaf_phicut_simple: the wedge between
   the half-planes at phi0 and phi1; a miss that exits at infinity flags -0 in y */
static void leaf_phicut(int* valid, f4* isect, const float* q0, float t_min, f3 o, f3 d)
{
    const float cosPhi0 = q0[0], sinPhi0 = q0[1], cosPhi1 = q0[2], sinPhi1 = q0[3];
    float d_n0 = d.x * sinPhi0 + d.y * (-cosPhi0);
    float d_n1 = d.x * (-sinPhi1) + d.y * (cosPhi1);
    float o_n0 = o.x * sinPhi0 + o.y * (-cosPhi0);
    float o_n1 = o.x * (-sinPhi1) + o.y * (cosPhi1);
    float t0 = d_n0 == 0.f ? t_min : -o_n0 / d_n0;
    float t1 = d_n1 == 0.f ? t_min : -o_n1 / d_n1;
    const float PR = d_n0 == 0.f ? -o_n0 : -d_n0;
    const float QR = d_n1 == 0.f ? -o_n1 : -d_n1;
    const float PQ = cosPhi0 * sinPhi1 - cosPhi1 * sinPhi0;
    int unbounded_exit = PQ >= 0.f ? (PR >= 0.f && QR <= 0.f) : (PR >= 0.f || QR <= 0.f);
    float side0 = o.x * cosPhi0 + o.y * sinPhi0 + (d.x * cosPhi0 + d.y * sinPhi0) * t0;
    float side1 = o.x * cosPhi1 + o.y * sinPhi1 + (d.x * cosPhi1 + d.y * sinPhi1) * t1;
    if (side0 < 0.f) t0 = t_min;
    if (side1 < 0.f) t1 = t_min;
    float t_near = fminf(t0, t1);
    float t_far = fmaxf(t0, t1);
    float t_cand = t_near > t_min ? t_near : (t_far > t_min ? t_far : t_min);
    *valid = t_cand > t_min;
    if (*valid) {
        isect->x = t_cand == t1 ? -sinPhi1 : sinPhi0;
        isect->y = t_cand == t1 ? cosPhi1 : -cosPhi0;
        isect->z = 0.f;
        isect->w = t_cand;
    } else if (unbounded_exit) {
        isect->y = -isect->y;
    }
}

/* csg_intersect_leaf_halfspace.h intersect_leaf_halfspace: n.x <= w side */
static void leaf_halfspace(int* valid, f4* isect, const float* q0, float t_min, f3 o, f3 d)
{
    f3 n = mk3(q0[0], q0[1], q0[2]);
    float w = q0[3];
    float on = dot3(o, n);
    float dn = dot3(d, n);
    float on_w = on - w;
    float adn = fabsf(dn);
    float eps = 1e-9f;
    int inside = on_w < -eps;
    float t = adn > 0.f ? -on_w / dn : t_min;
    *valid = t > t_min;
    if (*valid) {
        isect->x = n.x; isect->y = n.y; isect->z = n.z; isect->w = t;
    } else if (inside) {
        isect->y = -0.f;
    }
}

static int intersect_leaf(int* valid, f4* isect, const float* nd, const float* itra, const float* plan, float t_min,
                          f3 ro, f3 rd)
{
    *valid = 0;
    isect->x = 0.f; isect->y = 0.f; isect->z = 0.f; isect->w = 0.f;
    uint32_t typecode = nd_typecode(nd);
    uint32_t gt = nd_gtransformIdx(nd);
    int complement = nd_complement(nd);
    const float* q = gt > 0 ? itra + 16 * (gt - 1) : NULL;
    f3 origin = q ? q_right_multiply(q, ro, 1.f) : ro;
    f3 direction = q ? q_right_multiply(q, rd, 0.f) : rd;
    switch (typecode) {
        case CSG_SPHERE:   leaf_sphere(valid, isect, nd, t_min, origin, direction); break;
        case CSG_ZSPHERE:  leaf_zsphere(valid, isect, nd, nd + 4, t_min, origin, direction); break;
        case CSG_CYLINDER: leaf_cylinder(valid, isect, nd, nd + 4, t_min, origin, direction); break;
        case CSG_BOX3:     leaf_box3(valid, isect, nd, t_min, origin, direction); break;
        case CSG_CONE:     leaf_cone(valid, isect, nd, t_min, origin, direction); break;
        case CSG_CONVEXPOLYHEDRON: leaf_convexpolyhedron(valid, isect, nd, plan, t_min, origin, direction); break;
        case CSG_HYPERBOLOID: leaf_hyperboloid(valid, isect, nd, t_min, origin, direction); break;
        case CSG_PHICUT:   leaf_phicut(valid, isect, nd, t_min, origin, direction); break;
        case CSG_HALFSPACE: leaf_halfspace(valid, isect, nd, t_min, origin, direction); break;
        default: return -1;   /* leaf type not restated here */
    }
    if (*valid && q) q_left_multiply_inplace(q, isect, 0.f);
    if (complement) {
        isect->x = *valid ? -isect->x : -0.f;
        isect->y = *valid ? -isect->y : isect->y;
        isect->z = *valid ? -isect->z : isect->z;
    }
    return 0;
}

/* ------------------------------------------------------------ CSG tree */

/* Kensler single-hit CSG decision table, transcribed from the rules of
   "Ray Tracing CSG Objects Using Single Hit Intersections" as applied by
   CSG/csg_classify.h (LUT::lookup). act[op-1][closerB][stateA][stateB]. */
enum { A_MISS = 0, A_RETURN_A = 1, A_RETURN_B = 2, A_RETURN_FLIP_B = 3, A_LOOP_A = 4, A_LOOP_B = 5 };
enum { S_ENTER = 0, S_EXIT = 1, S_MISS = 2 };

static const unsigned char KENSLER[3][2][3][3] = {
    /* union */
    { /* A closer */ { { A_RETURN_A, A_LOOP_A, A_RETURN_A }, { A_RETURN_A, A_RETURN_B, A_RETURN_A }, { A_RETURN_B, A_RETURN_B, A_MISS } },
      /* B closer */ { { A_RETURN_B, A_RETURN_B, A_RETURN_A }, { A_LOOP_B, A_RETURN_A, A_RETURN_A }, { A_RETURN_B, A_RETURN_B, A_MISS } } },
    /* intersection */
    { { { A_LOOP_A, A_RETURN_A, A_MISS }, { A_LOOP_A, A_RETURN_A, A_MISS }, { A_MISS, A_MISS, A_MISS } },
      { { A_LOOP_B, A_LOOP_B, A_MISS }, { A_RETURN_B, A_RETURN_B, A_MISS }, { A_MISS, A_MISS, A_MISS } } },
    /* difference */
    { { { A_RETURN_A, A_LOOP_A, A_RETURN_A }, { A_RETURN_A, A_LOOP_A, A_RETURN_A }, { A_MISS, A_MISS, A_MISS } },
      { { A_LOOP_B, A_LOOP_B, A_RETURN_A }, { A_RETURN_FLIP_B, A_RETURN_FLIP_B, A_RETURN_A }, { A_MISS, A_MISS, A_MISS } } },
};

void oracle_packed_lut(uint32_t* acl, uint32_t* bcl)
{
    for (int op = 0; op < 3; ++op) {
        for (int cl = 0; cl < 2; ++cl) {
            uint32_t w = 0;
            for (int a = 0; a < 3; ++a)
                for (int b = 0; b < 3; ++b) {
                    unsigned off = 3u * a + b;
                    if (off < 8) w |= (uint32_t)KENSLER[op][cl][a][b] << (4 * off);
                }
            if (cl == 0) acl[op] = w; else bcl[op] = w;
        }
    }
}

static inline int csg_classify(f4 ise, f3 dir, float tmin)
{
    return fabsf(ise.w) > tmin ? (ise.x * dir.x + ise.y * dir.y + ise.z * dir.z < 0.f ? S_ENTER : S_EXIT) : S_MISS;
}

#define CSG_STACK_SIZE 15
#define TRANCHE_STACK_SIZE 4
#define ERR_OVERFLOW 0x400
#define ERR_TRANCHE_OVERFLOW 0x800
#define ERR_POP_EMPTY 0x1000
#define ERR_XOR_SIDE 0x2000
#define ERR_END_EMPTY 0x4000

static inline unsigned tree_depth(unsigned idx) { return 32u - (unsigned)__builtin_clz(idx) - 1u; }
static inline unsigned tree_height(unsigned n) { return (unsigned)(__builtin_ffs((int)n + 1) - 2); }
static inline unsigned postorder_next(unsigned i, unsigned elev) { return (i & 1u) ? i >> 1 : (i << elev) + (1u << elev); }
#define PACK_SLICE(b, e) ((((b) & 0xffu) << 16) | (((e) & 0xffu) << 24))

/* ------------------------------------------------------ CSG list nodes */
/* Compound nodes whose subNum leaves sit at root + subOffset + i (offsets from
   the prim's root node, CSG/CSGNode.h:187-211). */
#define LIST_PROPAGATE_EPSILON 0.0001f
#define LIST_MAX_ENTER 8

/* intersect_node_contiguous (CSG/csg_intersect_node.h:392-650): union of
   touching leaves. Zeroth pass: nearest enter and a count of exits; with no
   exit the ray starts outside and the nearest enter is the answer. Otherwise
   collect the enters, order them by t (indirect insertion sort), and march
   through them extending the farthest exit with each leaf entered before it. */
static int list_contiguous(int* valid, f4* isect, const float* nd, const float* root, const float* itra,
                           const float* plan, float t_min, f3 ro, f3 rd)
{
    const int num_sub = (int)nd_subNum(nd);
    const int off = (int)nd_subOffset(nd);
    f4 nearest_enter = mk4(0.f, 0.f, 0.f, RT_DEFAULT_MAX_F);
    f4 farthest_exit = mk4(0.f, 0.f, 0.f, t_min);
    f4 sub = mk4(0.f, 0.f, 0.f, 0.f);
    int exit_count = 0;
    for (int i = 0; i < num_sub; ++i) {
        int v = 0;
        if (intersect_leaf(&v, &sub, root + 16 * (off + i), itra, plan, t_min, ro, rd) != 0) return -1;
        if (!v) continue;
        int st = csg_classify(sub, rd, t_min);
        if (st == S_ENTER) { if (sub.w < nearest_enter.w) nearest_enter = sub; }
        else if (st == S_EXIT) exit_count += 1;
    }
    if (exit_count == 0) {
        *valid = nearest_enter.w > t_min && nearest_enter.w < RT_DEFAULT_MAX_F;
        if (*valid) *isect = nearest_enter;
        return 0;
    }
    int enter_count = 0;
    float enter[LIST_MAX_ENTER];
    int aux[LIST_MAX_ENTER], idx[LIST_MAX_ENTER];
    for (int isub = 0; isub < num_sub; ++isub) {
        int v = 0;
        if (intersect_leaf(&v, &sub, root + 16 * (off + isub), itra, plan, t_min, ro, rd) != 0) return -1;
        if (!v) continue;
        int st = csg_classify(sub, rd, t_min);
        if (st == S_ENTER) {
            if (enter_count >= LIST_MAX_ENTER) return -1;
            aux[enter_count] = isub; idx[enter_count] = enter_count; enter[enter_count] = sub.w;
            enter_count += 1;
        } else if (st == S_EXIT) {
            exit_count += 1;
            if (sub.w > farthest_exit.w) farthest_exit = sub;
        }
    }
    for (int i = 1; i < enter_count; ++i) {
        int key = idx[i];
        int j = i - 1;
        while (j >= 0 && enter[idx[j]] > enter[key]) { idx[j + 1] = idx[j]; j -= 1; }
        idx[j + 1] = key;
    }
    for (int i = 0; i < enter_count; ++i) {
        float tminAdvanced = enter[idx[i]] + LIST_PROPAGATE_EPSILON;
        int isub = aux[idx[i]];
        if (tminAdvanced < farthest_exit.w) {
            int v = 0;
            if (intersect_leaf(&v, &sub, root + 16 * (off + isub), itra, plan, tminAdvanced, ro, rd) != 0) return -1;
            if (v && csg_classify(sub, rd, tminAdvanced) == S_EXIT) {
                exit_count += 1;
                if (sub.w > farthest_exit.w) farthest_exit = sub;
            }
        }
    }
    *valid = exit_count > 0 && farthest_exit.w > t_min;
    if (*valid) *isect = farthest_exit;
    return 0;
}

/* intersect_node_discontiguous (CSG/csg_intersect_node.h:654-690): disjoint
   leaves, the closest valid leaf intersect */
static int list_discontiguous(int* valid, f4* isect, const float* nd, const float* root, const float* itra,
                              const float* plan, float t_min, f3 ro, f3 rd)
{
    const int num_sub = (int)nd_subNum(nd);
    const int off = (int)nd_subOffset(nd);
    f4 closest = mk4(0.f, 0.f, 0.f, RT_DEFAULT_MAX_F);
    f4 sub = mk4(0.f, 0.f, 0.f, 0.f);
    for (int i = 0; i < num_sub; ++i) {
        int v = 0;
        if (intersect_leaf(&v, &sub, root + 16 * (off + i), itra, plan, t_min, ro, rd) != 0) return -1;
        if (v && sub.w < closest.w) closest = sub;
    }
    *valid = closest.w < RT_DEFAULT_MAX_F;
    if (*valid) *isect = closest;
    return 0;
}

/* intersect_node_overlap (CSG/csg_intersect_node.h:819-905): intersection of
   leaves; each entered leaf is re-shot past its enter for its exit; a hit needs
   farthest enter < nearest exit with every leaf entered or exited */
static int list_overlap(int* valid, f4* isect, const float* nd, const float* root, const float* itra,
                        const float* plan, float t_min, f3 ro, f3 rd)
{
    const unsigned num_sub = nd_subNum(nd);
    const unsigned off = nd_subOffset(nd);
    f4 farthest_enter = mk4(0.f, 0.f, 0.f, t_min);
    f4 nearest_exit = mk4(0.f, 0.f, 0.f, RT_DEFAULT_MAX_F);
    f4 sub = mk4(0.f, 0.f, 0.f, 0.f);
    unsigned enter_count = 0, exit_count = 0;
    for (unsigned isub = 0; isub < num_sub; ++isub) {
        const float* sn = root + 16 * (off + isub);
        int v = 0;
        if (intersect_leaf(&v, &sub, sn, itra, plan, t_min, ro, rd) != 0) return -1;
        if (!v) continue;
        int st = csg_classify(sub, rd, t_min);
        if (st == S_ENTER) {
            enter_count += 1;
            if (sub.w > farthest_enter.w) farthest_enter = sub;
            float tminAdvanced = sub.w + LIST_PROPAGATE_EPSILON;
            if (intersect_leaf(&v, &sub, sn, itra, plan, tminAdvanced, ro, rd) != 0) return -1;
            if (v && csg_classify(sub, rd, tminAdvanced) == S_EXIT) {
                exit_count += 1;
                if (sub.w < nearest_exit.w) nearest_exit = sub;
            }
        } else if (st == S_EXIT) {
            exit_count += 1;
            if (sub.w < nearest_exit.w) nearest_exit = sub;
        }
    }
    *valid = 0;
    unsigned mx = enter_count > exit_count ? enter_count : exit_count;
    if (farthest_enter.w < nearest_exit.w && mx == num_sub) {
        if (farthest_enter.w > t_min && farthest_enter.w < RT_DEFAULT_MAX_F) { *valid = 1; *isect = farthest_enter; }
        else if (nearest_exit.w > t_min && nearest_exit.w < RT_DEFAULT_MAX_F) { *valid = 1; *isect = nearest_exit; }
    }
    return 0;
}

/* intersect_node (CSG/csg_intersect_node.h:913-937): list nodes or a leaf */
static int intersect_node(int* valid, f4* isect, const float* nd, const float* root, const float* itra,
                          const float* plan, float t_min, f3 ro, f3 rd)
{
    *valid = 0;
    switch (nd_typecode(nd)) {
        case CSG_CONTIGUOUS:    return list_contiguous(valid, isect, nd, root, itra, plan, t_min, ro, rd);
        case CSG_OVERLAP:       return list_overlap(valid, isect, nd, root, itra, plan, t_min, ro, rd);
        case CSG_DISCONTIGUOUS: return list_discontiguous(valid, isect, nd, root, itra, plan, t_min, ro, rd);
        default:                return intersect_leaf(valid, isect, nd, itra, plan, t_min, ro, rd);
    }
}

/* tree evaluations and loop decisions since the last oracle_tree_counts(reset);
   counted only once a reset has switched counting on (the shared counters would
   otherwise slow the multi-threaded runs that time the checker) */
static int g_tree_counting;
static int64_t g_tree_evals, g_tree_loops;

void oracle_tree_counts(int64_t* evals, int64_t* loops, int32_t reset)
{
    *evals = __atomic_load_n(&g_tree_evals, __ATOMIC_RELAXED);
    *loops = __atomic_load_n(&g_tree_loops, __ATOMIC_RELAXED);
    if (reset) {
        __atomic_store_n(&g_tree_evals, 0, __ATOMIC_RELAXED);
        __atomic_store_n(&g_tree_loops, 0, __ATOMIC_RELAXED);
        __atomic_store_n(&g_tree_counting, 1, __ATOMIC_RELAXED);
    }
}

/* intersect_tree (CSG/csg_intersect_tree.h:277-672) */
static int intersect_tree(int* valid, f4* isect, const float* node, const float* itra, const float* plan, float t_min,
                          f3 ro, f3 rd)
{
    const int counting = __atomic_load_n(&g_tree_counting, __ATOMIC_RELAXED);
    if (counting) __atomic_fetch_add(&g_tree_evals, 1, __ATOMIC_RELAXED);
    int numNode = (int)nd_subNum(node);
    unsigned height = tree_height((unsigned)numNode);
    const float propagate_epsilon = 0.0001f;
    int ierr = 0;
    float tr_tmin[TRANCHE_STACK_SIZE]; unsigned tr_slice[TRANCHE_STACK_SIZE]; int tr_curr = -1;
    f4 csg[CSG_STACK_SIZE]; int csg_curr = -1;

    unsigned fullTree = PACK_SLICE(1u << height, 0u);
    tr_curr++; tr_slice[tr_curr] = fullTree; tr_tmin[tr_curr] = t_min;

    while (tr_curr > -1) {
        unsigned slice = tr_slice[tr_curr]; float tmin = tr_tmin[tr_curr]; tr_curr--;
        unsigned beginIdx = (slice >> 16) & 0xffu;
        unsigned endIdx = (slice >> 24) & 0xffu;
        unsigned nodeIdx = beginIdx;
        while (nodeIdx != endIdx) {
            unsigned depth = tree_depth(nodeIdx);
            unsigned elevation = height - depth;
            const float* nd = node + 16 * (nodeIdx - 1);
            uint32_t typecode = nd_typecode(nd);
            if (typecode == CSG_ZERO) { nodeIdx = postorder_next(nodeIdx, elevation); continue; }
            int node_or_leaf = typecode >= CSG_NODE;
            if (node_or_leaf) {
                f4 nd_isect = mk4(0.f, 0.f, 0.f, 0.f);
                int nd_valid = 0;
                if (intersect_node(&nd_valid, &nd_isect, nd, node, itra, plan, tmin, ro, rd) != 0) return -1;
                nd_isect.w = copysignf(nd_isect.w, nodeIdx % 2 == 0 ? -1.f : 1.f);
                if (csg_curr >= CSG_STACK_SIZE - 1) { ierr = ERR_OVERFLOW; break; }
                csg_curr++; csg[csg_curr] = nd_isect;
            } else {
                if (csg_curr < 1) { ierr |= ERR_POP_EMPTY; break; }
                int firstLeft = signbit(csg[csg_curr].w) != 0;
                int secondLeft = signbit(csg[csg_curr - 1].w) != 0;
                if (!(firstLeft ^ secondLeft)) { ierr |= ERR_XOR_SIDE; break; }
                int left = firstLeft ? csg_curr : csg_curr - 1;
                int right = firstLeft ? csg_curr - 1 : csg_curr;
                int l_state = csg_classify(csg[left], rd, tmin);
                int r_state = csg_classify(csg[right], rd, tmin);
                float t_left = fabsf(csg[left].w);
                float t_right = fabsf(csg[right].w);
                int leftIsCloser = t_left <= t_right;
                int l_complement = signbit(csg[left].x) != 0;
                int r_complement = signbit(csg[right].x) != 0;
                int l_unbounded = signbit(csg[left].y) != 0;
                int r_unbounded = signbit(csg[right].y) != 0;
                int l_promote_miss = l_state == S_MISS && (l_complement || l_unbounded);
                int r_promote_miss = r_state == S_MISS && (r_complement || r_unbounded);
                if (r_promote_miss) { r_state = S_EXIT; leftIsCloser = 1; }
                if (l_promote_miss) { l_state = S_EXIT; leftIsCloser = 0; }
                int ctrl = (typecode >= CSG_UNION && typecode <= CSG_DIFFERENCE)
                         ? KENSLER[typecode - CSG_UNION][leftIsCloser ? 0 : 1][l_state][r_state] : A_MISS;
                if (ctrl < A_LOOP_A) {
                    f4 result = ctrl == A_MISS ? mk4(0.f, 0.f, 0.f, 0.f) : csg[ctrl == A_RETURN_A ? left : right];
                    if (ctrl == A_RETURN_FLIP_B) { result.x = -result.x; result.y = -result.y; result.z = -result.z; }
                    result.w = copysignf(result.w, nodeIdx % 2 == 0 ? -1.f : 1.f);
                    if (csg_curr < 0) { ierr = ERR_POP_EMPTY; break; } csg_curr--;
                    if (csg_curr < 0) { ierr = ERR_POP_EMPTY; break; } csg_curr--;
                    if (csg_curr >= CSG_STACK_SIZE - 1) { ierr = ERR_OVERFLOW; break; }
                    csg_curr++; csg[csg_curr] = result;
                } else {
                    if (counting) __atomic_fetch_add(&g_tree_loops, 1, __ATOMIC_RELAXED);
                    int loopside = ctrl == A_LOOP_A ? left : right;
                    int otherside = ctrl == A_LOOP_A ? right : left;
                    unsigned leftIdx = 2 * nodeIdx;
                    unsigned rightIdx = leftIdx + 1;
                    float tminAdvanced = fabsf(csg[loopside].w) + propagate_epsilon;
                    f4 other = csg[otherside];
                    if (csg_curr < 0) { ierr = ERR_POP_EMPTY; break; } csg_curr--;
                    if (csg_curr < 0) { ierr = ERR_POP_EMPTY; break; } csg_curr--;
                    if (csg_curr >= CSG_STACK_SIZE - 1) { ierr = ERR_OVERFLOW; break; }
                    csg_curr++; csg[csg_curr] = other;
                    unsigned endTree = PACK_SLICE(nodeIdx, endIdx);
                    unsigned leftTree = PACK_SLICE(leftIdx << (elevation - 1), rightIdx << (elevation - 1));
                    unsigned rightTree = PACK_SLICE(rightIdx << (elevation - 1), nodeIdx);
                    unsigned loopTree = ctrl == A_LOOP_A ? leftTree : rightTree;
                    if (tr_curr >= TRANCHE_STACK_SIZE - 1) { ierr = ERR_TRANCHE_OVERFLOW; break; }
                    tr_curr++; tr_slice[tr_curr] = endTree; tr_tmin[tr_curr] = tmin;
                    if (tr_curr >= TRANCHE_STACK_SIZE - 1) { ierr = ERR_TRANCHE_OVERFLOW; break; }
                    tr_curr++; tr_slice[tr_curr] = loopTree; tr_tmin[tr_curr] = tminAdvanced;
                    break;   /* act == BREAK: restart from the new tranche */
                }
            }
            nodeIdx = postorder_next(nodeIdx, elevation);
        }
        if (ierr) break;
    }
    ierr |= (csg_curr != 0) ? ERR_END_EMPTY : 0;
    if (csg_curr == 0) *isect = csg[0];
    *valid = isect->w > 0.f;
    (void)ierr;
    return 0;
}

/* intersect_prim (CSG/csg_intersect_tree.h:684-719) */
static int intersect_prim(int* valid, f4* isect, const float* node, const float* itra, const float* plan, float t_min,
                          f3 ro, f3 rd)
{
    uint32_t typecode = nd_typecode(node);
    *valid = 0;
    if (typecode >= CSG_LEAF) return intersect_leaf(valid, isect, node, itra, plan, t_min, ro, rd);
    if (typecode < CSG_NODE) return intersect_tree(valid, isect, node, itra, plan, t_min, ro, rd);
    if (typecode == CSG_CONTIGUOUS) return list_contiguous(valid, isect, node, node, itra, plan, t_min, ro, rd);
    if (typecode == CSG_DISCONTIGUOUS) return list_discontiguous(valid, isect, node, node, itra, plan, t_min, ro, rd);
    if (typecode == CSG_OVERLAP) return list_overlap(valid, isect, node, node, itra, plan, t_min, ro, rd);
    return -1;
}

int oracle_intersect_prim(const oracle_geom* g, int32_t prim_idx, float t_min,
                          const float* o, const float* d, float* isect4)
{
    if (!g || prim_idx < 0 || prim_idx >= g->n_prim) return -1;
    const int32_t* pr = (const int32_t*)(g->prim + 16 * prim_idx);
    const float* root = g->node + 16 * pr[1];
    f4 is = mk4(0.f, 0.f, 0.f, 0.f);
    int valid = 0;
    int rc = intersect_prim(&valid, &is, root, g->itra, g->plan, t_min, mk3(o[0], o[1], o[2]), mk3(d[0], d[1], d[2]));
    isect4[0] = is.x; isect4[1] = is.y; isect4[2] = is.z; isect4[3] = is.w;
    return rc != 0 ? rc : valid;
}

/* ------------------------------------------------------- world build */

/* Inverse of the affine instance transform in double precision, rounded to
   float once. Row-vector convention of qat4::right_multiply: rows 0..2 hold the
   inverse 3x3 (w column zero), row 3 the inverse translation. Cofactor formula. */
typedef struct { float m[16]; } w2o16;

static void invert_affine16(const float* m, float* out16)
{
    double a[3][3], t[3], inv[3][3];
    for (int r = 0; r < 3; ++r) for (int c = 0; c < 3; ++c) a[r][c] = (double)m[4 * r + c];
    for (int c = 0; c < 3; ++c) t[c] = (double)m[12 + c];
    double c00 = a[1][1] * a[2][2] - a[1][2] * a[2][1];
    double c01 = a[1][2] * a[2][0] - a[1][0] * a[2][2];
    double c02 = a[1][0] * a[2][1] - a[1][1] * a[2][0];
    double det = a[0][0] * c00 + a[0][1] * c01 + a[0][2] * c02;
    inv[0][0] = c00 / det;
    inv[0][1] = (a[0][2] * a[2][1] - a[0][1] * a[2][2]) / det;
    inv[0][2] = (a[0][1] * a[1][2] - a[0][2] * a[1][1]) / det;
    inv[1][0] = c01 / det;
    inv[1][1] = (a[0][0] * a[2][2] - a[0][2] * a[2][0]) / det;
    inv[1][2] = (a[0][2] * a[1][0] - a[0][0] * a[1][2]) / det;
    inv[2][0] = c02 / det;
    inv[2][1] = (a[0][1] * a[2][0] - a[0][0] * a[2][1]) / det;
    inv[2][2] = (a[0][0] * a[1][1] - a[0][1] * a[1][0]) / det;
    for (int r = 0; r < 3; ++r) {
        for (int c = 0; c < 3; ++c) out16[4 * r + c] = (float)inv[r][c];
        out16[4 * r + 3] = 0.f;
    }
    for (int c = 0; c < 3; ++c) out16[12 + c] = (float)(-(t[0] * inv[0][c] + t[1] * inv[1][c] + t[2] * inv[2][c]));
    out16[15] = 1.f;
}

static void box_empty(dbox* b) { for (int i = 0; i < 3; ++i) { b->lo[i] = 1e300; b->hi[i] = -1e300; } }
static void box_grow(dbox* b, const double* p) { for (int i = 0; i < 3; ++i) { if (p[i] < b->lo[i]) b->lo[i] = p[i]; if (p[i] > b->hi[i]) b->hi[i] = p[i]; } }
static void box_merge(dbox* b, const dbox* o) { for (int i = 0; i < 3; ++i) { if (o->lo[i] < b->lo[i]) b->lo[i] = o->lo[i]; if (o->hi[i] > b->hi[i]) b->hi[i] = o->hi[i]; } }

static int build_bvh(oworld* w, int first, int count, double (*cen)[3])
{
    int idx = w->n_nodes++;
    bnode* n = &w->nodes[idx];
    box_empty(&n->box);
    for (int i = first; i < first + count; ++i) box_merge(&n->box, &w->ibox[w->order[i]]);
    n->left = n->right = -1; n->first = first; n->count = count;
    if (count <= 4) return idx;
    dbox cb; box_empty(&cb);
    for (int i = first; i < first + count; ++i) box_grow(&cb, cen[w->order[i]]);
    int ax = 0; double ext = cb.hi[0] - cb.lo[0];
    for (int a = 1; a < 3; ++a) if (cb.hi[a] - cb.lo[a] > ext) { ext = cb.hi[a] - cb.lo[a]; ax = a; }
    /* median split by insertion-free partial ordering: simple sort on axis */
    for (int i = first + 1; i < first + count; ++i) {
        int v = w->order[i]; double key = cen[v][ax]; int j = i - 1;
        while (j >= first && cen[w->order[j]][ax] > key) { w->order[j + 1] = w->order[j]; --j; }
        w->order[j + 1] = v;
    }
    int half = count / 2;
    int l = build_bvh(w, first, half, cen);
    int r = build_bvh(w, first + half, count - half, cen);
    n = &w->nodes[idx];
    n->left = l; n->right = r; n->count = 0;
    return idx;
}

void* oracle_world_create(const oracle_geom* g)
{
    if (!g || g->n_inst <= 0) return NULL;
    oworld* w = (oworld*)calloc(1, sizeof(oworld));
    w->g = *g;
    int64_t ni = g->n_inst;
    w->w2o = calloc((size_t)ni, sizeof(w2o16));
    w->ibox = (dbox*)calloc((size_t)ni, sizeof(dbox));
    w->inst_gas = (int32_t*)calloc((size_t)ni, sizeof(int32_t));
    w->inst_ident = (uint32_t*)calloc((size_t)ni, sizeof(uint32_t));
    w->solid_np = (int32_t*)calloc((size_t)g->n_solid, sizeof(int32_t));
    w->solid_po = (int32_t*)calloc((size_t)g->n_solid, sizeof(int32_t));
    const unsigned char* sb = (const unsigned char*)g->solid;
    for (int64_t s = 0; s < g->n_solid; ++s) {
        memcpy(&w->solid_np[s], sb + 48 * s + 16, 4);
        memcpy(&w->solid_po[s], sb + 48 * s + 20, 4);
    }
    double (*cen)[3] = (double (*)[3])calloc((size_t)ni, sizeof(double[3]));
    for (int64_t i = 0; i < ni; ++i) {
        const float* m = g->inst + 16 * i;
        const int32_t* mi = (const int32_t*)m;
        w->inst_gas[i] = mi[7];            /* q1.i.w gas_idx            */
        w->inst_ident[i] = (uint32_t)mi[11];  /* q2.i.w sensor_identifier+1 */
        invert_affine16(m, ((w2o16*)w->w2o)[i].m);
        /* world box: transform the 8 corners of each prim box of the solid */
        int gas = w->inst_gas[i];
        box_empty(&w->ibox[i]);
        for (int p = 0; p < w->solid_np[gas]; ++p) {
            const float* pr = g->prim + 16 * (w->solid_po[gas] + p);
            double lo[3] = { pr[8], pr[9], pr[10] }, hi[3] = { pr[11], pr[12], pr[13] };
            for (int c = 0; c < 8; ++c) {
                double v[3] = { (c & 1) ? hi[0] : lo[0], (c & 2) ? hi[1] : lo[1], (c & 4) ? hi[2] : lo[2] };
                double pw[3];
                for (int k = 0; k < 3; ++k) pw[k] = v[0] * m[k] + v[1] * m[4 + k] + v[2] * m[8 + k] + m[12 + k];
                box_grow(&w->ibox[i], pw);
            }
        }
        for (int k = 0; k < 3; ++k) {
            double pad = 1e-3 * (w->ibox[i].hi[k] - w->ibox[i].lo[k]) + 1e-2;
            w->ibox[i].lo[k] -= pad; w->ibox[i].hi[k] += pad;
            cen[i][k] = 0.5 * (w->ibox[i].lo[k] + w->ibox[i].hi[k]);
        }
    }
    w->order = (int32_t*)calloc((size_t)ni, sizeof(int32_t));
    for (int64_t i = 0; i < ni; ++i) w->order[i] = (int32_t)i;
    w->nodes = (bnode*)calloc((size_t)(2 * ni + 1), sizeof(bnode));
    w->n_nodes = 0;
    build_bvh(w, 0, (int)ni, cen);
    free(cen);
    return w;
}

void oracle_world_destroy(void* wp)
{
    oworld* w = (oworld*)wp;
    if (!w) return;
    free(w->w2o); free(w->ibox); free(w->inst_gas); free(w->inst_ident);
    free(w->solid_np); free(w->solid_po); free(w->nodes); free(w->order);
    free(w);
}

static int ray_box(const dbox* b, f3 o, f3 d, double tmin, double tmax)
{
    double t0 = tmin, t1 = tmax;
    double oo[3] = { o.x, o.y, o.z }, dd[3] = { d.x, d.y, d.z };
    for (int k = 0; k < 3; ++k) {
        if (dd[k] == 0.0) { if (oo[k] < b->lo[k] || oo[k] > b->hi[k]) return 0; continue; }
        double inv = 1.0 / dd[k];
        double ta = (b->lo[k] - oo[k]) * inv, tb = (b->hi[k] - oo[k]) * inv;
        if (ta > tb) { double t = ta; ta = tb; tb = t; }
        if (ta > t0) t0 = ta;
        if (tb < t1) t1 = tb;
        if (t0 > t1) return 0;
    }
    return 1;
}

/* Object-frame ray for instance i: the world->object transform applied as
   qat4::right_multiply (w=1 for positions, w=0 terms dropped for directions). */
static inline void inst_ray(const float* m, f3 o, f3 d, f3* oo, f3* od)
{
    oo->x = m[0] * o.x + m[4] * o.y + m[8] * o.z + m[12];
    oo->y = m[1] * o.x + m[5] * o.y + m[9] * o.z + m[13];
    oo->z = m[2] * o.x + m[6] * o.y + m[10] * o.z + m[14];
    od->x = m[0] * d.x + m[4] * d.y + m[8] * d.z;
    od->y = m[1] * d.x + m[5] * d.y + m[9] * d.z;
    od->z = m[2] * d.x + m[6] * d.y + m[10] * d.z;
}

/* Candidate test of all prims of instance i: IS program of CSGOptiX7.cu:896-967
   with closest-hit acceptance t < best (ties broken on the smaller
   (instance, prim) so the result does not depend on traversal order). */
static int test_instance(const oworld* w, int64_t i, f3 o, f3 d, float tmin, float* best_t,
                         uint64_t* best_key, oprd* prd)
{
    const float* m = ((const w2o16*)w->w2o)[i].m;
    f3 oo, od;
    inst_ray(m, o, d, &oo, &od);
    int gas = w->inst_gas[i];
    for (int p = 0; p < w->solid_np[gas]; ++p) {
        int64_t pidx = w->solid_po[gas] + p;
        const int32_t* pr = (const int32_t*)(w->g.prim + 16 * pidx);
        const float* root = w->g.node + 16 * pr[1];
        f4 is = mk4(0.f, 0.f, 0.f, 0.f);
        int valid = 0;
        if (intersect_prim(&valid, &is, root, w->g.itra, w->g.plan, tmin, oo, od) != 0) return -1;
        if (!valid) continue;
        if (!(is.w >= tmin)) continue;   /* outside [tmin, tmax]: ignored like optixReportIntersection */
        uint64_t key = ((uint64_t)i << 32) | (uint64_t)p;
        if (is.w < *best_t || (is.w == *best_t && key < *best_key)) {
            *best_t = is.w; *best_key = key;
            /* closest-hit: object normal to world (transform by the transpose of
               the world->object linear part), CSGOptiX7.cu:843-853 */
            f3 n;
            n.x = m[0] * is.x + m[1] * is.y + m[2] * is.z;
            n.y = m[4] * is.x + m[5] * is.y + m[6] * is.z;
            n.z = m[8] * is.x + m[9] * is.y + m[10] * is.z;
            prd->normal = n;
            prd->t = is.w;
            f3 lpos = add3(oo, smul3(is.w, od));
            prd->lposcost = lpos.z / sqrtf(dot3(lpos, lpos));
            prd->lposfphi = 0.f;
            prd->lx = lpos.x;
            prd->ly = lpos.y;
            uint32_t boundary = nd_boundary(root);
            uint32_t gp = (uint32_t)((const int32_t*)(w->g.prim + 16 * pidx))[15];
            prd->gp_boundary = ((gp & 0xffffu) << 16) | (boundary & 0xffffu);
            prd->iindex_identity = (((uint32_t)i & 0xffffu) << 16) | (w->inst_ident[i] & 0xffffu);
            prd->inst = (uint32_t)i;
        }
    }
    return 0;
}

/* trace (CSGOptiX7.cu:113-218) + miss (CSGOptiX7.cu:682-709) */
static int world_trace(const oworld* w, f3 o, f3 d, float tmin, float tmax, oprd* prd, int traverse)
{
    float best = tmax;
    uint64_t best_key = ~0ull;
    prd->normal = mk3(0.f, 0.f, 0.f); prd->t = 0.f;
    prd->lposcost = 0.f; prd->lposfphi = 0.f;
    prd->lx = 0.f; prd->ly = NAN;
    prd->iindex_identity = 0xffffffffu; prd->gp_boundary = 0xffffffffu; prd->inst = 0xffffffffu;
    if (traverse == 0) {
        for (int64_t i = 0; i < w->g.n_inst; ++i) {
            if (!ray_box(&w->ibox[i], o, d, (double)tmin, (double)best)) continue;
            if (test_instance(w, i, o, d, tmin, &best, &best_key, prd) != 0) return -1;
        }
    } else {
        int stack[128]; int sp = 0;
        stack[sp++] = 0;
        while (sp > 0) {
            const bnode* n = &w->nodes[stack[--sp]];
            if (!ray_box(&n->box, o, d, (double)tmin, (double)best)) continue;
            if (n->count > 0) {
                for (int k = n->first; k < n->first + n->count; ++k) {
                    int64_t i = w->order[k];
                    if (!ray_box(&w->ibox[i], o, d, (double)tmin, (double)best)) continue;
                    if (test_instance(w, i, o, d, tmin, &best, &best_key, prd) != 0) return -1;
                }
            } else {
                if (sp + 2 > 128) return -2;
                stack[sp++] = n->left; stack[sp++] = n->right;
            }
        }
    }
    return 0;
}

int oracle_trace(void* wp, const float* rays, int64_t n, float* isect, int32_t traverse)
{
    oworld* w = (oworld*)wp;
    int bad = 0;
#pragma omp parallel for schedule(dynamic, 256) reduction(|:bad)
    for (int64_t i = 0; i < n; ++i) {
        const float* r = rays + 8 * i;
        oprd prd;
        if (world_trace(w, mk3(r[0], r[1], r[2]), mk3(r[4], r[5], r[6]), r[3], r[7], &prd, traverse) != 0) bad |= 1;
        float* out = isect + 8 * i;
        int miss = (prd.gp_boundary & 0xffffu) == 0xffffu;
        f3 nn = miss ? prd.normal : norm3(prd.normal);
        out[0] = nn.x; out[1] = nn.y; out[2] = nn.z; out[3] = prd.t;
        uint32_t* ou32 = (uint32_t*)out;
        ou32[4] = prd.iindex_identity;
        ou32[5] = prd.gp_boundary & 0xffffu;
        memcpy(&out[6], &prd.lposcost, 4);
        ou32[7] = prd.inst;
    }
    return bad ? -1 : 0;
}

/* ------------------------------------------------------------ physics */

typedef struct {
    f4 material1, m1group2, material2, surface;
    int32_t optical[4];
    int32_t index[4];
    int m1_line;
} ostate;

/*
 * Adjoint score accumulator (no counterpart in the reference). For the
 * segment of a photon inside material m with absorption length La and
 * Rayleigh length Ls the path-probability factor is exp(-ell/La) exp(-ell/Ls),
 * times 1/La when the segment ends in absorption (reemission) or 1/Ls when it
 * ends in a Rayleigh scatter. Its log-derivatives
 *   d/dLa = ell/La^2 - [absorbed]/La,   d/dLs = ell/Ls^2 - [scattered]/Ls
 * are spread onto the parameter bins of the two table samples the FP32 lerp
 * of bnd_lookup reads, with weights (1-a) and a. row layout per parameter
 * block: [n_bin absorption bins][n_bin Rayleigh bins].
 */
typedef struct {
    const int32_t* mat_block;   /* (n_mat+1) block of each 1-based material index, <0 none */
    int32_t n_mat;
    const int32_t* sample_bin;  /* (n_wl) bin of each wavelength sample */
    int32_t n_bin;
    float* row;                 /* (n_block*2*n_bin) accumulated in path order */
} oscore;

enum { SEG_SAIL = 0, SEG_ABSORB = 1, SEG_SCATTER = 2 };

static void score_seg(oscore* sc, const oracle_phys* ph, int m1_line, float wl, float La, float Ls, float ell, int ev)
{
    if (!sc) return;
    int mat = ph->optical[4 * m1_line];
    if (mat < 0 || mat > sc->n_mat) return;
    int blk = sc->mat_block[mat];
    if (blk < 0) return;
    int nwl = ph->n_wl;
    float fx = (wl - ph->wl_low) / ph->wl_step;
    float fl = floorf(fx);
    float a = fx - fl;
    int i0 = (int)fl;
    i0 = ((i0 % nwl) + nwl) % nwl;
    int i1 = (i0 + 1) % nwl;
    int b0 = sc->sample_bin[i0], b1 = sc->sample_bin[i1];
    float sa = ell / (La * La) - (ev == SEG_ABSORB ? 1.f / La : 0.f);
    float ss = ell / (Ls * Ls) - (ev == SEG_SCATTER ? 1.f / Ls : 0.f);
    float* r = sc->row + blk * 2 * sc->n_bin;
    r[b0] += sa * (1.f - a);
    r[b1] += sa * a;
    r[sc->n_bin + b0] += ss * (1.f - a);
    r[sc->n_bin + b1] += ss * a;
}

/* Linear interpolation of one bnd "texture" row at fractional texel fx,
   wrap addressing (qudarap/QTex.cc:293-305, qudarap/qbnd.h:108-133). */
static f4 bnd_lookup(const oracle_phys* ph, float nm, int line, int k)
{
    int nwl = ph->n_wl;
    float fx = (nm - ph->wl_low) / ph->wl_step;
    float fl = floorf(fx);
    float a = fx - fl;
    int i0 = (int)fl;
    i0 = ((i0 % nwl) + nwl) % nwl;
    int i1 = (i0 + 1) % nwl;
    int64_t iy = 2 * (int64_t)line + k;
    const float* row = ph->bnd + iy * (int64_t)nwl * 4;
    const float* v0 = row + 4 * i0;
    const float* v1 = row + 4 * i1;
    return mk4(v0[0] + a * (v1[0] - v0[0]), v0[1] + a * (v1[1] - v0[1]),
               v0[2] + a * (v1[2] - v0[2]), v0[3] + a * (v1[3] - v0[3]));
}

/* qbnd::fill_state (qudarap/qbnd.h:192-223) */
static void fill_state(const oracle_phys* ph, ostate* s, uint32_t boundary, float wavelength, float cosTheta)
{
    int line = (int)boundary * 4;
    int m1_line = cosTheta > 0.f ? line + 3 : line + 0;
    int m2_line = cosTheta > 0.f ? line + 0 : line + 3;
    int su_line = cosTheta > 0.f ? line + 2 : line + 1;
    s->material1 = bnd_lookup(ph, wavelength, m1_line, 0);
    s->m1group2 = bnd_lookup(ph, wavelength, m1_line, 1);
    s->material2 = bnd_lookup(ph, wavelength, m2_line, 0);
    s->surface = bnd_lookup(ph, wavelength, su_line, 0);
    for (int c = 0; c < 4; ++c) s->optical[c] = ph->optical[4 * su_line + c];
    s->index[0] = ph->optical[4 * m1_line];
    s->index[1] = ph->optical[4 * m2_line];
    s->index[2] = ph->optical[4 * su_line];
    s->index[3] = 0;
    s->m1_line = m1_line;
}

/* scintillation ICDF "texture": qscint::wavelength_hd0/10/20 (qudarap/qscint.h:160-230),
   linear filtering with normalized coordinates and wrap addressing */
static float icdf_tex(const oracle_phys* ph, float x, int layer)
{
    int n = ph->n_icdf;
    float xb = x * (float)n - 0.5f;
    float fl = floorf(xb);
    float a = xb - fl;
    int i0 = (int)fl;
    i0 = ((i0 % n) + n) % n;
    int i1 = (i0 + 1) % n;
    int L = layer < ph->n_icdf_layer ? layer : 0;
    const float* row = ph->icdf + (int64_t)L * n;
    return row[i0] + a * (row[i1] - row[i0]);
}

static float scint_wavelength(const oracle_phys* ph, float u0)
{
    switch (ph->icdf_hd_factor) {
        case 0: return icdf_tex(ph, u0, 0);
        case 10: {
            int lhs = u0 < 0.1f, rhs = u0 > 0.9f;
            float x = lhs ? (u0 * 10.f) : (rhs ? (u0 - 0.9f) * 10.f : u0);
            return icdf_tex(ph, x, lhs ? 1 : (rhs ? 2 : 0));
        }
        case 20: {
            int lhs = u0 < 0.05f, rhs = u0 > 0.95f;
            float x = lhs ? (u0 * 20.f) : (rhs ? (u0 - 0.95f) * 20.f : u0);
            return icdf_tex(ph, x, lhs ? 1 : (rhs ? 2 : 0));
        }
        default: return 0.f;
    }
}

/* smath::rotateUz (sysrap/smath.h:95-115) */
static void rotateUz(f3* d, f3 u)
{
    float up = u.x * u.x + u.y * u.y;
    if (up > 0.f) {
        up = sqrtf(up);
        float px = d->x, py = d->y, pz = d->z;
        d->x = (u.x * u.z * px - u.y * py) / up + u.x * pz;
        d->y = (u.y * u.z * px + u.x * py) / up + u.y * pz;
        d->z = -up * px + u.z * pz;
    } else if (u.z < 0.f) {
        d->x = -d->x;
        d->z = -d->z;
    }
}

/* qsim::uniform_sphere(u0,u1) (qudarap/qsim.h:180-186) */
static f3 uniform_sphere(float u0, float u1)
{
    float phi = u0 * 2.f * PI_F;
    float cosTheta = 2.f * u1 - 1.f;
    float sinTheta = sqrtf(1.f - cosTheta * cosTheta);
    float sp, cp;
    oa_sincosf(phi, &sp, &cp);
    return mk3(cp * sinTheta, sp * sinTheta, cosTheta);
}

/* qsim::random_direction_marsaglia (qudarap/qsim.h:554-575) */
static void random_direction_marsaglia(f3* dir, orng* rng)
{
    float u, v, b, a;
    int trials = 0;
    do {
        float u0 = ou(rng);
        float u1 = ou(rng);
        u = 2.f * u0 - 1.f;
        v = 2.f * u1 - 1.f;
        b = u * u + v * v;
    } while (b > 1.f && ++trials < OA_MAX_TRIALS);
    if (b > 1.f) { u = 0.f; v = 0.f; b = 0.f; }   /* bound reached: (0, 0, -1) */
    a = 2.f * sqrtf(1.f - b);
    dir->x = a * u;
    dir->y = a * v;
    dir->z = 2.f * b - 1.f;
}

/* qsim::lambertian_direction (qudarap/qsim.h:433-486) */
static void lambertian_direction(f3* dir, f3 normal, float orient, orng* rng)
{
    float ndotv;
    int count = 0;
    float u;
    do {
        count++;
        random_direction_marsaglia(dir, rng);
        ndotv = dot3(*dir, normal) * orient;
        if (ndotv < 0.f) {
            *dir = smul3(-1.f, *dir);
            ndotv = -1.f * ndotv;
        }
        u = ou(rng);
    } while (!(u < ndotv) && (count < 1024));
}

/* qsim::rayleigh_scatter (qudarap/qsim.h:604-692) */
static void rayleigh_scatter(oph* p, orng* rng)
{
    f3 direction, polarization;
    int looping = 1;
    int trials = 0;
    do {
        float u0 = ou(rng), u1 = ou(rng), u2 = ou(rng), u3 = ou(rng), u4 = ou(rng);
        float cosTheta = u0;
        float sinTheta = sqrtf(1.0f - u0 * u0);
        if (u1 < 0.5f) cosTheta = -cosTheta;
        float sinPhi, cosPhi;
        oa_sincosf(2.f * PI_F * u2, &sinPhi, &cosPhi);
        direction.x = sinTheta * cosPhi;
        direction.y = sinTheta * sinPhi;
        direction.z = cosTheta;
        rotateUz(&direction, p->mom);
        float constant = -dot3(direction, p->pol);
        polarization.x = p->pol.x + constant * direction.x;
        polarization.y = p->pol.y + constant * direction.y;
        polarization.z = p->pol.z + constant * direction.z;
        if (dot3(polarization, polarization) == 0.f) {
            oa_sincosf(2.f * PI_F * u3, &sinPhi, &cosPhi);
            polarization.x = cosPhi;
            polarization.y = sinPhi;
            polarization.z = 0.f;
            rotateUz(&polarization, direction);
        } else {
            if (u3 < 0.5f) polarization = neg3(polarization);
        }
        polarization = norm3(polarization);
        float doCosTheta = dot3(polarization, p->pol);
        float doCosTheta2 = doCosTheta * doCosTheta;
        looping = doCosTheta2 < u4;
    } while (looping && ++trials < OA_MAX_TRIALS);
    p->mom = direction;
    p->pol = polarization;
}

/* qsim::propagate_to_boundary (qudarap/qsim.h:721-866) */
static int propagate_to_boundary(uint32_t* flag, orng* rng, oph* p, const ostate* s, float distance_to_boundary,
                                 const oracle_phys* ph, oscore* sc)
{
    float absorption_length = s->material1.y;
    float scattering_length = s->material1.z;
    float reemission_prob = s->material1.w;
    float group_velocity = s->m1group2.x;
    float u_scattering = ou(rng);
    float u_absorption = ou(rng);
    float scattering_distance = -scattering_length * oa_logf(u_scattering);
    float absorption_distance = -absorption_length * oa_logf(u_absorption);
    if (absorption_distance <= scattering_distance) {
        if (absorption_distance <= distance_to_boundary) {
            score_seg(sc, ph, s->m1_line, p->wavelength, absorption_length, scattering_length, absorption_distance,
                      SEG_ABSORB);
            p->time += absorption_distance / group_velocity;
            p->pos = add3(p->pos, smul3(absorption_distance, p->mom));
            float u_reemit = reemission_prob == 0.f ? 2.f : ou(rng);
            if (u_reemit < reemission_prob) {
                float u_re_wavelength = ou(rng);
                float u_re_mom_ph = ou(rng);
                float u_re_mom_ct = ou(rng);
                float u_re_pol_ph = ou(rng);
                float u_re_pol_ct = ou(rng);
                p->wavelength = scint_wavelength(ph, u_re_wavelength);
                p->mom = uniform_sphere(u_re_mom_ph, u_re_mom_ct);
                p->pol = norm3(cross3(uniform_sphere(u_re_pol_ph, u_re_pol_ct), p->mom));
                *flag = F_BULK_REEMIT;
                return C_CONTINUE;
            } else {
                *flag = F_BULK_ABSORB;
                return C_BREAK;
            }
        }
    } else {
        if (scattering_distance <= distance_to_boundary) {
            score_seg(sc, ph, s->m1_line, p->wavelength, absorption_length, scattering_length, scattering_distance,
                      SEG_SCATTER);
            p->time += scattering_distance / group_velocity;
            p->pos = add3(p->pos, smul3(scattering_distance, p->mom));
            rayleigh_scatter(p, rng);
            *flag = F_BULK_SCATTER;
            return C_CONTINUE;
        }
    }
    score_seg(sc, ph, s->m1_line, p->wavelength, absorption_length, scattering_length, distance_to_boundary, SEG_SAIL);
    p->pos = add3(p->pos, smul3(distance_to_boundary, p->mom));
    p->time += distance_to_boundary / group_velocity;
    return C_BOUNDARY;
}

static inline float len2(float x, float y) { return sqrtf(x * x + y * y); }

/* qsim::propagate_at_boundary (qudarap/qsim.h:1001-1208) */
static int propagate_at_boundary_core(uint32_t* flag, float u_reflect, oph* p, float n1, float n2, f3 normal,
                                      float* transcoeff)
{
    float eta = n1 / n2;
    float _c1 = -dot3(p->mom, normal);
    f3 oriented_normal = _c1 < 0.f ? neg3(normal) : normal;
    f3 trans = cross3(p->mom, oriented_normal);
    float trans_length = len3(trans);
    int normal_incidence = trans_length < 1e-6f;
    float inv_tl = 1.0f / trans_length;   /* float3/float is a*(1/s) in sysrap/scuda.h:561-565 */
    f3 A_trans = normal_incidence ? p->pol : mul3s(trans, inv_tl);
    float E1_perp = dot3(p->pol, A_trans);
    float c1 = fabsf(_c1);
    float c2c2 = 1.f - eta * eta * (1.f - c1 * c1);
    int tir = c2c2 < 0.f;
    float EdotN = dot3(p->pol, oriented_normal);
    float c2 = tir ? 0.f : sqrtf(c2c2);
    float n1c1 = n1 * c1;
    float n2c2 = n2 * c2;
    float n2c1 = n2 * c1;
    float n1c2 = n1 * c2;
    float E1x, E1y;
    if (normal_incidence) { E1x = 0.f; E1y = 1.f; }
    else { E1x = E1_perp; E1y = len3(sub3(p->pol, smul3(E1_perp, A_trans))); }
    float E2_tx = 2.f * n1c1 * E1x / (n1c1 + n2c2);
    float E2_ty = 2.f * n1c1 * E1y / (n2c1 + n1c2);
    float E2_rx = E2_tx - E1x;
    float E2_ry = (n2 * E2_ty / n1) - E1y;
    float invR = 1.0f / sqrtf(E2_rx * E2_rx + E2_ry * E2_ry);
    float RRx = E2_rx * invR, RRy = E2_ry * invR;
    float invT = 1.0f / sqrtf(E2_tx * E2_tx + E2_ty * E2_ty);
    float TTx = E2_tx * invT, TTy = E2_ty * invT;
    float TransCoeff = (tir || n1c1 == 0.f) ? 0.f : n2c2 * (E2_tx * E2_tx + E2_ty * E2_ty) / n1c1;
    if (transcoeff) *transcoeff = TransCoeff;
    int reflect = u_reflect > TransCoeff;
    p->mom = reflect ? add3(p->mom, smul3(2.0f * c1, oriented_normal))
                     : add3(smul3(eta, p->mom), smul3(eta * c1 - c2, oriented_normal));
    f3 A_paral = norm3(cross3(p->mom, A_trans));
    if (normal_incidence) {
        p->pol = reflect ? mul3s(p->pol, n2 > n1 ? -1.f : 1.f) : p->pol;
    } else if (reflect) {
        p->pol = tir ? add3(neg3(p->pol), smul3(2.f * EdotN, oriented_normal))
                     : add3(smul3(RRx, A_trans), smul3(RRy, A_paral));
    } else {
        p->pol = add3(smul3(TTx, A_trans), smul3(TTy, A_paral));
    }
    *flag = reflect ? F_BOUNDARY_REFLECT : F_BOUNDARY_TRANSMIT;
    return C_CONTINUE;
}

/* qsim::reflect_diffuse (qudarap/qsim.h:1988-2023) */
static void reflect_diffuse(oph* p, f3 normal, orng* rng)
{
    f3 old_mom = p->mom;
    float orient = dot3(old_mom, normal) > 0.f ? -1.f : 1.f;
    lambertian_direction(&p->mom, normal, orient, rng);
    f3 facet_normal = norm3(sub3(p->mom, old_mom));
    float EdotN = dot3(p->pol, facet_normal);
    p->pol = add3(smul3(-1.f, p->pol), smul3(2.f * EdotN, facet_normal));
}

/* qsim::reflect_specular (qudarap/qsim.h:2041-2093, not WITH_ORIENT) */
static void reflect_specular(oph* p, f3 normal)
{
    float PdotN = dot3(p->mom, normal);
    p->mom = sub3(p->mom, smul3(2.f * PdotN, normal));
    float EdotN = dot3(p->pol, normal);
    p->pol = add3(smul3(-1.f, p->pol), smul3(2.f * EdotN, normal));
}

/* qsim::propagate_at_surface (qudarap/qsim.h:1680-1750, not WITH_CUSTOM4) */
static int propagate_at_surface(uint32_t* flag, orng* rng, oph* p, const ostate* s, f3 normal)
{
    float detect = s->surface.x;
    float absorb = s->surface.y;
    float reflect_diffuse_ = s->surface.w;
    float u_surface = ou(rng);
    int action = u_surface < absorb + detect ? C_BREAK : C_CONTINUE;
    if (action == C_BREAK) {
        *flag = u_surface < absorb ? F_SURFACE_ABSORB : F_SURFACE_DETECT;
    } else {
        *flag = u_surface < absorb + detect + reflect_diffuse_ ? F_SURFACE_DREFLECT : F_SURFACE_SREFLECT;
        if (*flag == F_SURFACE_DREFLECT) reflect_diffuse(p, normal, rng);
        else reflect_specular(p, normal);
    }
    return action;
}

/* qsim::propagate (qudarap/qsim.h:2229-2338) */
static int propagate(orng* rng, oph* p, const oprd* prd, const oracle_phys* ph, oscore* sc)
{
    uint32_t boundary = prd->gp_boundary & 0xffffu;
    uint32_t identity = prd->iindex_identity & 0xffffu;
    uint32_t iindex = prd->iindex_identity >> 16;
    float lposcost = prd->lposcost;
    f3 normal = prd->normal;
    float cosTheta = dot3(p->mom, normal);
    ph_set_prd(p, boundary, identity, cosTheta, iindex);
    ostate s;
    fill_state(ph, &s, boundary, p->wavelength, cosTheta);
    uint32_t flag = 0;
    int command = propagate_to_boundary(&flag, rng, p, &s, prd->t, ph, sc);
    if (command == C_BOUNDARY) {
        int ems = s.optical[1];
        if (ems == EMS_NOSURFACE) {
            float u_reflect = ou(rng);
            command = propagate_at_boundary_core(&flag, u_reflect, p, s.material1.x, s.material2.x, normal, NULL);
        } else if (ems == EMS_SURFACE) {
            command = propagate_at_surface(&flag, rng, p, &s, normal);
        } else if (lposcost < 0.f) {
            command = propagate_at_surface(&flag, rng, p, &s, normal);
        } else if (ems == EMS_ZPLUS_SENSOR_A) {
            (void)ou(rng);   /* u_surface_burn, qsim.h:1753-1758 */
            flag = F_SURFACE_DETECT;
            command = C_BREAK;
        }
    }
    ph_set_flag(p, flag);
    return command;
}

/* ------------------------------------------------------- generation */

/* storch::generate (sysrap/storch.h:192-520) */
static void torch_generate(oph* p, orng* rng, const float* gs, uint64_t photon_id)
{
    const uint32_t* gu = (const uint32_t*)gs;
    f3 gpos = mk3(gs[4], gs[5], gs[6]);
    float gtime = gs[7];
    f3 gmom = mk3(gs[8], gs[9], gs[10]);
    float gweight = gs[11];
    float gwl = gs[15];
    float zen_x = gs[16], zen_y = gs[17], azi_x = gs[18], azi_y = gs[19];
    float radius = gs[20], distance = gs[21];
    uint32_t numphoton = gu[3];
    uint32_t type = gu[23];
    memset(p, 0, sizeof(*p));
    if (type == T_DISC) {
        p->wavelength = gwl; p->time = gtime; p->mom = gmom;
        float u_zenith = zen_x + ou(rng) * (zen_y - zen_x);
        float u_azimuth = azi_x + ou(rng) * (azi_y - azi_x);
        float r = radius * u_zenith;
        float phi = 2.f * PI_F * u_azimuth;
        float sinPhi, cosPhi; oa_sincosf(phi, &sinPhi, &cosPhi);
        p->pos = mk3(r * cosPhi, r * sinPhi, 0.f);
        rotateUz(&p->pos, p->mom);
        p->pos = add3(p->pos, gpos);
        p->pol = mk3(sinPhi, -cosPhi, 0.f);
        rotateUz(&p->pol, p->mom);
    } else if (type == T_SPHERE || type == T_POINT_ISO) {
        p->wavelength = gwl; p->time = gtime;
        float u_zenith = zen_x + ou(rng) * (zen_y - zen_x);
        float u_azimuth = azi_x + ou(rng) * (azi_y - azi_x);
        float phi = 2.f * PI_F * u_azimuth;
        float sinPhi, cosPhi; oa_sincosf(phi, &sinPhi, &cosPhi);
        float cosTheta = 1.f - 2.0f * u_zenith;
        float sinTheta = sqrtf(1.0f - cosTheta * cosTheta);
        float flip = copysignf(1.f, radius);
        p->mom = mk3(flip * sinTheta * cosPhi, flip * sinTheta * sinPhi, flip * cosTheta);
        float rad = fabsf(radius);
        p->pos = mk3(sinTheta * cosPhi * rad, sinTheta * sinPhi * rad, cosTheta * rad);
        float frac_twopi = distance;
        float phase = 2.f * PI_F * frac_twopi;
        float sph, cph; oa_sincosf(phase, &sph, &cph);
        p->pol = mk3(cph, sph, 0.f);
        rotateUz(&p->pol, p->mom);
        if (type == T_POINT_ISO) {
            /* this project's isotropic point source: T_SPHERE about gs.pos,
               wavelength uniform in [gs.wavelength, gs.wavelength+gs.weight] */
            p->pos = add3(p->pos, gpos);
            if (gweight > 0.f) p->wavelength = gwl + ou(rng) * gweight;
        }
    } else if (type == T_SPHERE_MARSAGLIA) {
        p->wavelength = gwl; p->time = gtime;
        float u, v, b, a;
        int trials = 0;
        do {
            float u0z = zen_x + ou(rng) * (zen_y - zen_x);
            float u1a = azi_x + ou(rng) * (azi_y - azi_x);
            u = 2.f * u0z - 1.f; v = 2.f * u1a - 1.f; b = u * u + v * v;
        } while (b > 1.f && ++trials < OA_MAX_TRIALS);
        if (b > 1.f) { u = 0.f; v = 0.f; b = 0.f; }
        a = 2.f * sqrtf(1.f - b);
        float rad = fabsf(radius);
        float flip = copysignf(1.f, radius);
        p->mom = mk3(flip * a * u, flip * a * v, flip * (2.f * b - 1.f));
        p->pos = mk3(a * u * rad, a * v * rad, (2.f * b - 1.f) * rad);
        float phase = 2.f * PI_F * distance;
        float sph, cph; oa_sincosf(phase, &sph, &cph);
        p->pol = mk3(cph, sph, 0.f);
        rotateUz(&p->pol, p->mom);
    } else if (type == T_LINE) {
        p->wavelength = gwl; p->time = gtime; p->mom = gmom;
        float frac = (float)photon_id / (float)numphoton;
        float sfrac = 2.f * (frac - 0.5f);
        float r = radius * sfrac;
        p->pos = mk3(r, 0.f, 0.f);
        rotateUz(&p->pos, p->mom);
        p->pos = add3(p->pos, gpos);
        p->pol = mk3(0.f, -1.f, 0.f);
        rotateUz(&p->pol, p->mom);
    } else if (type == T_POINT) {
        p->wavelength = gwl; p->time = gtime; p->mom = gmom;
        p->pos = mk3(0.f, 0.f, 0.f);
        rotateUz(&p->pos, p->mom);
        p->pos = add3(p->pos, gpos);
        p->pol = mk3(0.f, -1.f, 0.f);
        rotateUz(&p->pol, p->mom);
    } else if (type == T_CIRCLE) {
        p->wavelength = gwl; p->time = gtime;
        float f = (float)photon_id / (float)numphoton;
        float frac = azi_x * (1.f - f) + azi_y * (f);
        float phi = 2.f * PI_F * frac;
        float sinPhi, cosPhi; oa_sincosf(phi, &sinPhi, &cosPhi);
        float r = radius < 0.f ? -radius : radius;
        p->mom = mk3(radius < 0.f ? -cosPhi : cosPhi, 0.f, radius < 0.f ? -sinPhi : sinPhi);
        p->pos = mk3(r * cosPhi, 0.f, r * sinPhi);
        p->pos = add3(p->pos, gpos);
        p->pol = mk3(0.f, -1.f, 0.f);
        rotateUz(&p->pol, p->mom);
    } else if (type == T_RECTANGLE) {
        /* storch.h:463-516: four sides (left, right, bottom, top) of the rectangle
           x in [azimuth.x, azimuth.y], z in [zenith.x, zenith.y], momenta inwards;
           photons past 4*(numphoton/4) keep zero position offset and momentum;
           with fewer than 4 photons (a zero side size) all go to side 0 */
        p->wavelength = gwl; p->time = gtime;
        int side_size = (int)(numphoton / 4u);
        uint64_t ss = side_size > 0 ? (uint64_t)side_size : 1u;
        int side = (int)(photon_id / ss);
        int side_index = (int)(photon_id - (uint64_t)side * ss);
        float frac = (float)side_index / (float)(side_size > 0 ? side_size : 1);
        if (side == 0 || side == 1) {
            p->pos = mk3(side == 0 ? azi_x : azi_y, 0.f, (1.f - frac) * zen_x + frac * zen_y);
            p->mom = mk3(side == 0 ? 1.f : -1.f, 0.f, 0.f);
        } else if (side == 2 || side == 3) {
            p->pos = mk3((1.f - frac) * azi_x + frac * azi_y, 0.f, side == 2 ? zen_x : zen_y);
            p->mom = mk3(0.f, 0.f, side == 2 ? 1.f : -1.f);
        }
        p->pos = add3(p->pos, gpos);
        p->pol = mk3(0.f, -1.f, 0.f);
        rotateUz(&p->pol, p->mom);
    }
    ph_zero_flags(p);
    ph_set_flag(p, F_TORCH);
}

/* qscint::generate + mom_pol_wavelength (qudarap/qscint.h:66-158) */
static void scint_generate(oph* p, orng* rng, const float* gs, const oracle_phys* ph)
{
    memset(p, 0, sizeof(*p));
    float u0 = ou(rng), u1 = ou(rng), u2 = ou(rng), u3 = ou(rng);
    float cost = 1.f - 2.f * u0;
    float sint = sqrtf((1.f - cost) * (1.f + cost));
    float phi = 2.f * PI_F * u1;
    float sinp, cosp; oa_sincosf(phi, &sinp, &cosp);
    p->mom = mk3(sint * cosp, sint * sinp, cost);
    p->pol = mk3(cost * cosp, cost * sinp, -sint);
    phi = 2.f * PI_F * u2;
    oa_sincosf(phi, &sinp, &cosp);
    p->pol = norm3(add3(smul3(cosp, p->pol), smul3(sinp, cross3(p->mom, p->pol))));
    p->wavelength = scint_wavelength(ph, u3);
    f3 gpos = mk3(gs[4], gs[5], gs[6]);
    float gtime = gs[7];
    f3 dpos = mk3(gs[8], gs[9], gs[10]);
    float step_length = gs[11];
    float charge = gs[13];
    float meanVelocity = gs[15];
    float ScintillationTime = gs[20];
    float fraction = charge == 0.f ? 1.f : ou(rng);
    p->pos = add3(gpos, smul3(fraction, dpos));
    float u4 = ou(rng);
    float deltaTime = fraction * step_length / meanVelocity - ScintillationTime * oa_logf(u4);
    p->time = gtime + deltaTime;
    ph_zero_flags(p);
    ph_set_flag(p, F_SCINTILLATION);
}

/* qcerenkov::generate (qudarap/qcerenkov.h:generate, wavelength_sampled_bndtex,
   fraction_sampled) on the scerenkov genstep layout (sysrap/scerenkov.h:29-58):
   q1 pos,time  q2 DeltaPosition,step_length  q3 code,charge,weight,preVelocity
   q4 BetaInverse,Wmin,Wmax,maxCos  q5 maxSin2,MeanNumberOfPhotons1,2,postVelocity */
static void cerenkov_generate(oph* p, orng* rng, const float* gs, const oracle_phys* ph)
{
    const int matline = ((const int32_t*)gs)[2];
    const f3 gpos = mk3(gs[4], gs[5], gs[6]);
    const float gtime = gs[7];
    const f3 dpos = mk3(gs[8], gs[9], gs[10]);
    const float step_length = gs[11];
    const float preVelocity = gs[15];
    const float BetaInverse = gs[16], Wmin = gs[17], Wmax = gs[18];
    const float maxSin2 = gs[20], N1 = gs[21], N2 = gs[22], postVelocity = gs[23];
    memset(p, 0, sizeof(*p));
    const f3 p0 = norm3(dpos);
    /* wavelength_sampled_bndtex: reciprocal sampling flat in energy, rejection on sin^2 */
    float wavelength, cosTheta, sin2Theta;
    unsigned count = 0;
    float u_maxSin2;
    do {
        float u0 = ou(rng);
        float w = Wmin + u0 * (Wmax - Wmin);
        wavelength = Wmin * Wmax / w;
        f4 props = bnd_lookup(ph, wavelength, matline, 0);
        float sampledRI = props.x;
        cosTheta = BetaInverse / sampledRI;
        sin2Theta = fmaxf(0.f, (1.f - cosTheta) * (1.f + cosTheta));
        float u1 = ou(rng);
        u_maxSin2 = u1 * maxSin2;
        count += 1;
    } while (u_maxSin2 > sin2Theta && count < 100);
    float sinTheta = sqrtf(sin2Theta);
    float u0 = ou(rng);
    float phi = 2.f * PI_F * u0;
    float sinPhi, cosPhi;
    oa_sincosf(phi, &sinPhi, &cosPhi);
    p->mom = mk3(sinTheta * cosPhi, sinTheta * sinPhi, cosTheta);
    rotateUz(&p->mom, p0);
    p->pol = mk3(cosTheta * cosPhi, cosTheta * sinPhi, -sinTheta);
    rotateUz(&p->pol, p0);
    p->wavelength = wavelength;
    /* fraction_sampled: position along the step with the linear N1 -> N2 yield */
    float fraction, delta, NumberOfPhotons, N, u;
    float MeanNumberOfPhotonsMax = fmaxf(N1, N2);
    float DeltaN = (N1 - N2);
    int trials = 0;
    do {
        fraction = ou(rng);
        delta = fraction * step_length;
        NumberOfPhotons = N1 - fraction * DeltaN;
        u = ou(rng);
        N = u * MeanNumberOfPhotonsMax;
    } while (N > NumberOfPhotons && ++trials < OA_MAX_TRIALS);
    float midVelocity = preVelocity + fraction * (postVelocity - preVelocity) * 0.5f;
    p->time = gtime + delta / midVelocity;
    p->pos = add3(gpos, smul3(fraction, dpos));
    ph_zero_flags(p);
    ph_set_flag(p, F_CERENKOV);
}

/* ---------------------------------------------------------- simulate */

static inline void seq_add_nibble(uint64_t* seqhis, uint64_t* seqbnd, unsigned slot, unsigned flag, unsigned boundary)
{
    unsigned iseq = slot / 16u;
    unsigned shift = 4u * (slot - iseq * 16u);
    if (iseq < 2u) {
        unsigned ffs = flag ? (unsigned)__builtin_ffs((int)flag) : 0u;
        seqhis[iseq] |= ((uint64_t)(ffs & 0xfu)) << shift;
        seqbnd[iseq] |= ((uint64_t)(boundary & 0xfu)) << shift;
    }
}

/* genstep of in-event photon idx: off[g] <= idx < off[g+1] (qudarap/QEvt.cu:181-226 seeding) */
static int64_t genstep_of(const int64_t* off, int64_t n_gs, int64_t idx)
{
    int64_t lo = 0, hi = n_gs;
    while (hi - lo > 1) { int64_t mid = (lo + hi) / 2; if (off[mid] <= idx) lo = mid; else hi = mid; }
    while (lo < n_gs && off[lo + 1] <= idx) ++lo;
    return lo;
}

/* one photon through the CSGOptiX7.cu:432-530 simulate body; returns error bits */
/* words skipped at the start of every photon stream of an event */
static uint64_t event_skip(const oracle_cfg* cfg, int32_t event_idx)
{
    return cfg->rng_offset + cfg->event_skipahead * (uint64_t)event_idx;
}

/* xorwow jump matrix of an event (NULL when not xorwow or no skip); free() it */
static uint32_t* xorwow_jump(const oracle_cfg* cfg, int32_t event_idx)
{
    const uint64_t n = event_skip(cfg, event_idx);
    if (cfg->rng_kind != 2 || n == 0) return NULL;
    uint32_t* J = (uint32_t*)malloc(sizeof(uint32_t) * OA_XORWOW_JUMP_WORDS);
    uint32_t* work = (uint32_t*)malloc(sizeof(uint32_t) * OA_XORWOW_WORK_WORDS);
    oa_xorwow_jump_matrix(n, J, work);
    free(work);
    return J;
}

static int simulate_one(const oworld* w, const oracle_phys* ph, const oracle_cfg* cfg, int32_t event_idx,
                        const float* gs, uint64_t photon_idx, oph* p, uint64_t* seqhis, uint64_t* seqbnd, oscore* sc,
                        oprd* last, const uint32_t* xwJ)
{
    int bad = 0;
    orng rng;
    rng.kind = cfg->rng_kind;
    if (last) { memset(last, 0, sizeof(*last)); last->ly = NAN; }
    if (rng.kind == 1) {
        /* every photon needs its row: the runtime rejects shorter arrays */
        if ((int64_t)photon_idx >= cfg->precooked_ni || cfg->precooked_nk <= 0) { memset(p, 0, sizeof(*p)); return 8; }
        rng.nk = cfg->precooked_nk;
        rng.cursor = 0;
        rng.pre = cfg->precooked + (int64_t)photon_idx * cfg->precooked_nk;
    } else if (rng.kind == 2) {
        oa_xorwow_seed(&rng.xw, cfg->seed, photon_idx);
        oa_xorwow_skip(&rng.xw, xwJ, event_skip(cfg, event_idx));
    } else {
        oa_rng_init(&rng.ph, cfg->seed, cfg->rng_offset, cfg->event_skipahead, (uint64_t)event_idx, photon_idx);
    }
    int gencode = ((const int32_t*)gs)[0];
    if (gencode == GS_TORCH) torch_generate(p, &rng, gs, photon_idx);
    else if (gencode == GS_SCINTILLATION || gencode == GS_SCINT_R4695) scint_generate(p, &rng, gs, ph);
    else if (gencode == GS_CERENKOV || gencode == GS_G4CERENKOV_MODIFIED) cerenkov_generate(p, &rng, gs, ph);
    else { memset(p, 0, sizeof(*p)); bad |= 2; }
    ph_set_index(p, photon_idx);
    seqhis[0] = seqhis[1] = seqbnd[0] = seqbnd[1] = 0;
    int bounce = 0;
    seq_add_nibble(seqhis, seqbnd, (unsigned)bounce, ph_flag(p), ph_boundary(p));
    oprd prd;
    memset(&prd, 0, sizeof(prd));
    prd.ly = NAN;
    while (bounce < cfg->max_bounce && p->time < cfg->max_time) {
        float tmin = (p->orient_boundary_flag & cfg->eps0_mask) ? cfg->tmin0 : cfg->tmin;
        if (world_trace(w, p->pos, p->mom, tmin, cfg->tmax, &prd, cfg->traverse) != 0) { bad |= 1; break; }
        if ((prd.gp_boundary & 0xffffu) == 0xffffu) break;
        prd.normal = norm3(prd.normal);
        int command = propagate(&rng, p, &prd, ph, sc);
        bounce++;
        seq_add_nibble(seqhis, seqbnd, (unsigned)bounce, ph_flag(p), ph_boundary(p));
        if (command == C_BREAK) break;
    }
    if (last) *last = prd;
    return bad;
}

/* sphotonlite::init + set_lpos (sysrap/sphotonlite.h:192-198,234-244) of the
   final photon and its last trace (CSGOptiX7.cu:520-526), float->u16 clamped */
static void make_lite(const oph* p, const oprd* prd, uint32_t* out4)
{
    float fphi = (prd->ly != prd->ly) ? 0.f : oa_fphi(prd->ly, prd->lx);
    float qc = fminf(fmaxf(prd->lposcost * 65535.f + 0.5f, 0.f), 65535.f);
    float qf = fminf(fmaxf(fphi * 65535.f + 0.5f, 0.f), 65535.f);
    float t = p->time;
    uint32_t tu;
    memcpy(&tu, &t, 4);
    out4[0] = (1u << 16) | (p->identity & 0xffffu);
    out4[1] = tu;
    out4[2] = ((uint32_t)qc << 16) | (uint32_t)qf;
    out4[3] = p->flagmask;
}

static int64_t* genstep_offsets(const float* gensteps, int64_t n_gs)
{
    int64_t* off = (int64_t*)malloc(sizeof(int64_t) * (size_t)(n_gs + 1));
    off[0] = 0;
    for (int64_t g = 0; g < n_gs; ++g) off[g + 1] = off[g] + (int64_t)((const uint32_t*)gensteps)[24 * g + 3];
    return off;
}

int64_t oracle_simulate_ex(void* wp, const oracle_phys* ph, const oracle_cfg* cfg, int32_t event_idx,
                           const float* gensteps, int64_t n_gs, int64_t photon_offset,
                           int64_t p_begin, int64_t p_end, float* photon_out, uint64_t* seq_out, uint32_t* lite_out)
{
    oworld* w = (oworld*)wp;
    if (!w || !ph || !cfg || (n_gs > 0 && !gensteps)) return -1;
    int64_t* off = genstep_offsets(gensteps, n_gs);
    int64_t total = off[n_gs];
    if (p_end > total) p_end = total;
    if (p_begin < 0) p_begin = 0;
    int bad = 0;
    uint32_t* xwJ = xorwow_jump(cfg, event_idx);
#ifdef _OPENMP
    if (cfg->nthreads > 0) omp_set_num_threads(cfg->nthreads);
#endif
#pragma omp parallel for schedule(dynamic, 256) reduction(|:bad)
    for (int64_t idx = p_begin; idx < p_end; ++idx) {
        const float* gs = gensteps + 24 * genstep_of(off, n_gs, idx);
        oph p;
        uint64_t seqhis[2], seqbnd[2];
        oprd last;
        bad |= simulate_one(w, ph, cfg, event_idx, gs, (uint64_t)(photon_offset + idx), &p, seqhis, seqbnd, NULL, &last,
                            xwJ);
        int64_t o = idx - p_begin;
        if (photon_out) memcpy(photon_out + 16 * o, &p, 64);
        if (lite_out) make_lite(&p, &last, lite_out + 4 * o);
        if (seq_out) {
            seq_out[4 * o + 0] = seqhis[0]; seq_out[4 * o + 1] = seqhis[1];
            seq_out[4 * o + 2] = seqbnd[0]; seq_out[4 * o + 3] = seqbnd[1];
        }
    }
    free(xwJ);
    free(off);
    return bad ? -(int64_t)bad - 1 : total;
}

int64_t oracle_simulate(void* wp, const oracle_phys* ph, const oracle_cfg* cfg, int32_t event_idx,
                        const float* gensteps, int64_t n_gs, int64_t photon_offset,
                        int64_t p_begin, int64_t p_end, float* photon_out, uint64_t* seq_out)
{
    return oracle_simulate_ex(wp, ph, cfg, event_idx, gensteps, n_gs, photon_offset, p_begin, p_end, photon_out,
                              seq_out, NULL);
}

float oracle_atan2f(float y, float x) { return oa_atan2f(y, x); }

void oracle_atan2f_n(const float* y, const float* x, float* out, int64_t n)
{
    for (int64_t i = 0; i < n; ++i) out[i] = oa_atan2f(y[i], x[i]);
}

int oracle_adjoint_rows(void* wp, const oracle_phys* ph, const oracle_cfg* cfg, int32_t event_idx,
                        const float* gensteps, int64_t n_gs, int64_t photon_offset,
                        const int64_t* photons, int64_t n, const int32_t* mat_block, int32_t n_mat,
                        const int32_t* sample_bin, int32_t n_bin, int32_t n_param, float* rows)
{
    oworld* w = (oworld*)wp;
    if (!w || !ph || !cfg || !gensteps || n_gs <= 0 || (n > 0 && (!photons || !rows))) return -1;
    int64_t* off = genstep_offsets(gensteps, n_gs);
    const int64_t total = off[n_gs];
    int bad = 0;
    uint32_t* xwJ = xorwow_jump(cfg, event_idx);
#ifdef _OPENMP
    if (cfg->nthreads > 0) omp_set_num_threads(cfg->nthreads);
#endif
#pragma omp parallel for schedule(dynamic, 64) reduction(|:bad)
    for (int64_t i = 0; i < n; ++i) {
        float* row = rows + i * (int64_t)n_param;
        for (int32_t k = 0; k < n_param; ++k) row[k] = 0.f;
        const int64_t idx = photons[i];
        if (idx < 0 || idx >= total) { bad |= 4; continue; }
        oscore sc;
        sc.mat_block = mat_block; sc.n_mat = n_mat; sc.sample_bin = sample_bin; sc.n_bin = n_bin; sc.row = row;
        oph p;
        uint64_t seqhis[2], seqbnd[2];
        bad |= simulate_one(w, ph, cfg, event_idx, gensteps + 24 * genstep_of(off, n_gs, idx),
                            (uint64_t)(photon_offset + idx), &p, seqhis, seqbnd, &sc, NULL, xwJ);
    }
    free(xwJ);
    free(off);
    return bad ? -bad : 0;
}

/*
 * Simtrace: qsim::generate_photon_simtrace_frame (qudarap/qsim.h:2460-2522) for
 * OpticksGenstep_FRAME gensteps (q0 = gencode, gridaxes, gsid, numphoton; q1 =
 * local position; q2..q5 = the frame transform as a qat4), the trace of
 * CSGOptiX7.cu:563-604 on RNG stream (event 0, photon index), and the record
 * of sevent::add_simtrace (sysrap/sevent.h:670-696):
 *   q0 hit normal (world, not normalized) and t   q1 pos + t dir, tmin
 *   q2 start position, globalPrimIdx<<16|boundary  q3 direction, iindex<<16|identity
 * A miss writes q0 = (0.6,0.6,0.6,1): the miss program's background colour and
 * t = 1 (CSGOptiX7.cu:683-700), so q1 = pos + dir.
 */
enum { GRID_XYZ = 0, GRID_YZ = 1, GRID_XZ = 2, GRID_XY = 3 };

static void simtrace_photon(const oworld* w, const oracle_cfg* cfg, const float* gs, uint64_t photon_idx, float* rec)
{
    orng rng;
    rng.kind = 0;
    oa_rng_init(&rng.ph, cfg->seed, cfg->rng_offset, cfg->event_skipahead, 0, photon_idx);
    const int gridaxes = ((const int32_t*)gs)[1];
    const float* qt = gs + 8;                     /* q2..q5 */
    float u0 = ou(&rng);
    float sinPhi, cosPhi;
    oa_sincosf(2.f * PI_F * u0, &sinPhi, &cosPhi);
    float u1 = ou(&rng);
    float cosTheta = 2.f * u1 - 1.f;
    float sinTheta = sqrtf(1.f - cosTheta * cosTheta);
    f3 dir;
    switch (gridaxes) {
        case GRID_YZ: dir = mk3(0.f, cosPhi, sinPhi); break;
        case GRID_XZ: dir = mk3(cosPhi, 0.f, sinPhi); break;
        case GRID_XY: dir = mk3(cosPhi, sinPhi, 0.f); break;
        default:      dir = mk3(sinTheta * cosPhi, sinTheta * sinPhi, cosTheta); break;
    }
    f3 pos = q_right_multiply(qt, mk3(gs[4], gs[5], gs[6]), 1.f);
    dir = q_right_multiply(qt, dir, 0.f);
    oprd prd;
    world_trace(w, pos, dir, cfg->tmin, cfg->tmax, &prd, cfg->traverse);
    const int miss = (prd.gp_boundary & 0xffffu) == 0xffffu;
    const float t = miss ? 1.f : prd.t;
    /* a miss carries the miss program's background colour, midgrey (CSGOptiX/SBT.cc:185-191) */
    rec[0] = miss ? 0.6f : prd.normal.x; rec[1] = miss ? 0.6f : prd.normal.y; rec[2] = miss ? 0.6f : prd.normal.z; rec[3] = t;
    rec[4] = pos.x + t * dir.x; rec[5] = pos.y + t * dir.y; rec[6] = pos.z + t * dir.z; rec[7] = cfg->tmin;
    rec[8] = pos.x; rec[9] = pos.y; rec[10] = pos.z;
    uint32_t gpb = miss ? 0xffffffffu : prd.gp_boundary, ii = miss ? 0xffffffffu : prd.iindex_identity;
    memcpy(rec + 11, &gpb, 4);
    rec[12] = dir.x; rec[13] = dir.y; rec[14] = dir.z;
    memcpy(rec + 15, &ii, 4);
}

int64_t oracle_simtrace(void* wp, const oracle_cfg* cfg, const float* gensteps, int64_t n_gs, float* out)
{
    oworld* w = (oworld*)wp;
    if (!w || !cfg || n_gs <= 0 || !gensteps || !out) return -1;
    int64_t* off = genstep_offsets(gensteps, n_gs);
    const int64_t total = off[n_gs];
#ifdef _OPENMP
    if (cfg->nthreads > 0) omp_set_num_threads(cfg->nthreads);
#endif
#pragma omp parallel for schedule(dynamic, 256)
    for (int64_t idx = 0; idx < total; ++idx)
        simtrace_photon(w, cfg, gensteps + 24 * genstep_of(off, n_gs, idx), (uint64_t)idx, out + 16 * idx);
    free(off);
    return total;
}

/* ------------------------------------------------------- test helpers */

float oracle_logf(float x) { return oa_logf(x); }
void oracle_sincosf(float x, float* s, float* c) { oa_sincosf(x, s, c); }

void oracle_philox_block(const uint32_t* ctr4, uint32_t k0, uint32_t k1, uint32_t* out4)
{
    oa_u32x4 c; for (int i = 0; i < 4; ++i) c.v[i] = ctr4[i];
    oa_u32x4 o = oa_philox4x32_10(c, k0, k1);
    for (int i = 0; i < 4; ++i) out4[i] = o.v[i];
}

void oracle_uniforms(uint64_t seed, uint64_t offset, uint64_t skip, uint64_t event_idx, uint64_t photon,
                     int32_t n, float* out)
{
    oa_philox_stream s;
    oa_rng_init(&s, seed, offset, skip, event_idx, photon);
    for (int i = 0; i < n; ++i) out[i] = oa_rng_uniform(&s);
}

void oracle_xorwow_words(uint64_t seed, uint64_t offset, uint64_t skip, uint64_t event_idx, uint64_t photon,
                         int32_t n, uint32_t* out, uint32_t* state6)
{
    oracle_cfg c;
    memset(&c, 0, sizeof(c));
    c.rng_kind = 2; c.rng_offset = offset; c.event_skipahead = skip;
    uint32_t* J = xorwow_jump(&c, (int32_t)event_idx);
    oa_xorwow_state s;
    oa_xorwow_seed(&s, seed, photon);
    oa_xorwow_skip(&s, J, event_skip(&c, (int32_t)event_idx));
    if (state6) { state6[0] = s.x0; state6[1] = s.x1; state6[2] = s.x2; state6[3] = s.x3; state6[4] = s.x4; state6[5] = s.d; }
    for (int i = 0; i < n; ++i) out[i] = oa_xorwow_next(&s);
    free(J);
}

void oracle_xorwow_from_state(const uint32_t* state6, uint64_t skip, int32_t n, uint32_t* out, uint32_t* state_out)
{
    oa_xorwow_state s;
    s.x0 = state6[0]; s.x1 = state6[1]; s.x2 = state6[2]; s.x3 = state6[3]; s.x4 = state6[4]; s.d = state6[5];
    uint32_t* J = NULL;
    if (skip) {
        J = (uint32_t*)malloc(sizeof(uint32_t) * OA_XORWOW_JUMP_WORDS);
        uint32_t* work = (uint32_t*)malloc(sizeof(uint32_t) * OA_XORWOW_WORK_WORDS);
        oa_xorwow_jump_matrix(skip, J, work);
        free(work);
    }
    oa_xorwow_skip(&s, J, skip);
    if (state_out) {
        state_out[0] = s.x0; state_out[1] = s.x1; state_out[2] = s.x2; state_out[3] = s.x3; state_out[4] = s.x4;
        state_out[5] = s.d;
    }
    for (int i = 0; i < n; ++i) out[i] = oa_xorwow_next(&s);
    free(J);
}

void oracle_xorwow_jump(uint64_t n, uint32_t* J800)
{
    uint32_t* work = (uint32_t*)malloc(sizeof(uint32_t) * OA_XORWOW_WORK_WORDS);
    oa_xorwow_jump_matrix(n, J800, work);
    free(work);
}

int oracle_step_boundary(const float* mom, const float* pol, const float* normal, float n1, float n2,
                         float u_reflect, float* mom_out, float* pol_out, uint32_t* flag_out, float* tc_out)
{
    oph p; memset(&p, 0, sizeof(p));
    p.mom = mk3(mom[0], mom[1], mom[2]);
    p.pol = mk3(pol[0], pol[1], pol[2]);
    uint32_t flag = 0;
    propagate_at_boundary_core(&flag, u_reflect, &p, n1, n2, mk3(normal[0], normal[1], normal[2]), tc_out);
    mom_out[0] = p.mom.x; mom_out[1] = p.mom.y; mom_out[2] = p.mom.z;
    pol_out[0] = p.pol.x; pol_out[1] = p.pol.y; pol_out[2] = p.pol.z;
    *flag_out = flag;
    return 0;
}

void oracle_rayleigh(const float* mom, const float* pol, uint64_t seed, uint64_t photon, int32_t n,
                     float* mom_out, float* pol_out)
{
    for (int i = 0; i < n; ++i) {
        oph p; memset(&p, 0, sizeof(p));
        p.mom = mk3(mom[0], mom[1], mom[2]);
        p.pol = mk3(pol[0], pol[1], pol[2]);
        orng r; r.kind = 0;
        oa_rng_init(&r.ph, seed, 0, 0, 0, photon + (uint64_t)i);
        rayleigh_scatter(&p, &r);
        mom_out[3 * i + 0] = p.mom.x; mom_out[3 * i + 1] = p.mom.y; mom_out[3 * i + 2] = p.mom.z;
        pol_out[3 * i + 0] = p.pol.x; pol_out[3 * i + 1] = p.pol.y; pol_out[3 * i + 2] = p.pol.z;
    }
}

void oracle_logf_n(const float* x, float* y, int64_t n)
{
    for (int64_t i = 0; i < n; ++i) y[i] = oa_logf(x[i]);
}

void oracle_sincosf_n(const float* x, float* s, float* c, int64_t n)
{
    for (int64_t i = 0; i < n; ++i) oa_sincosf(x[i], &s[i], &c[i]);
}
