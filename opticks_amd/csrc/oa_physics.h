/*
 * oa_physics.h : qsim propagation physics, photon generation and world trace
 * for gfx950. Function-by-function counterparts:
 *   fill                 qbnd::fill_state            qudarap/qbnd.h:192-223
 *   to_boundary          qsim::propagate_to_boundary qudarap/qsim.h:721-866
 *   rayleigh             qsim::rayleigh_scatter      qudarap/qsim.h:604-692
 *   at_boundary          qsim::propagate_at_boundary qudarap/qsim.h:1001-1208
 *   at_surface           qsim::propagate_at_surface  qudarap/qsim.h:1680-1758
 *   reflect_diffuse/specular, lambertian, marsaglia  qudarap/qsim.h:433-575,1988-2093
 *   propagate            qsim::propagate             qudarap/qsim.h:2229-2338
 *   gen_torch/gen_scint  storch::generate / qscint::generate
 *                        sysrap/storch.h:192-520, qudarap/qscint.h:66-200
 *   trace                optixTrace + IS/CH/MS       CSGOptiX/CSGOptiX7.cu:113-218,682-967
 */
#pragma once
#include "oa_device.h"

namespace oa {

/* ----------------------------------------------------------- tables */

/* Wrap addressing of a table of n texels: i0 = i mod n in [0, n) and its
   neighbour (i0 + 1) mod n. The integer divisions run only for an index
   outside the table (wavelengths beyond the domain), the results are the same. */
__device__ static __attribute__((noinline)) int wrap_index(int i, int n) { return ((i % n) + n) % n; }

/* OOL: the integer divisions out of line (a smaller kernel; the taped forward's register
   allocation does better with them inline: wrap_ool) */
template <bool OOL = true>
OA_D void wrap_pair(int i, int n, int& i0, int& i1)
{
    i0 = i;
    if (i0 < 0 || i0 >= n) i0 = OOL ? wrap_index(i0, n) : ((i0 % n) + n) % n;
    i1 = i0 + 1 == n ? 0 : i0 + 1;
}

/* One lerped float4 of bnd row (2*line+k) at wavelength nm: the FP32 form of
   the linearly filtered, wrap-addressed texture fetch of qbnd::boundary_lookup. */
template <bool OOL = true>
OA_D float4 bnd_lookup(const SimParams& P, float nm, int line, int k)
{
    const int nwl = P.n_wl;
    float fx = (nm - P.wl_low) / P.wl_step;
    float fl = floorf(fx);
    float a = fx - fl;
    int i0, i1;
    wrap_pair<OOL>((int)fl, nwl, i0, i1);
    const float4* row = P.bnd + (int64_t)(2 * line + k) * nwl;
    float4 v0 = row[i0];
    float4 v1 = row[i1];
    return make_float4(v0.x + a * (v1.x - v0.x), v0.y + a * (v1.y - v0.y),
                       v0.z + a * (v1.z - v0.z), v0.w + a * (v1.w - v0.w));
}

struct State {
    float4 material1;     /* n, absorption, scattering, reemission */
    float  group_velocity;
    float  n2;
    float4 surface;       /* detect, absorb, specular, diffuse */
    int    ems;
    int    m1_line;       /* bnd line of the material the photon travels in */
};

/* Segment events reported to a Tape during propagate_to_boundary. */
enum : int { SEG_SAIL = 0, SEG_ABSORB = 1, SEG_SCATTER = 2 };

/* The forward kernels record nothing. Every propagation step makes exactly one
   segment call (propagate -> to_boundary), so a photon's step number is its
   segment number: `at` names it before the step, `finish` gets the count. */
struct NoTape {
    __device__ inline __attribute__((always_inline)) void at(const SimParams&, int64_t, int) {}
    __device__ inline __attribute__((always_inline)) void seg(const SimParams&, int, float, float, float, float, int) {}
    __device__ inline __attribute__((always_inline)) void finish(const SimParams&, int64_t, int) {}
};

/* A forward kernel run for a following adjoint keeps, per photon, the arguments
   of its segment calls -- bnd line, segment event, wavelength and segment length
   (the absorption and scattering lengths are the table values at that line and
   wavelength, looked up again when the record is read) -- up to P.tape_k of them,
   and the number it had. The adjoint then scores a hit from its records instead
   of replaying the photon. It adds no per-lane state: it holds the photon index
   and step number the lane keeps anyway, and forms the record address from them
   and the launch fields only at the store. */
struct RecordTape {
    int64_t g;
    int32_t step;
    __device__ inline __attribute__((always_inline)) void at(const SimParams&, int64_t g_, int step_)
    {
        g = g_; step = step_;
    }
    __device__ inline __attribute__((always_inline)) void seg(const SimParams& P, int line, float wl, float, float, float ell,
                                                              int ev)
    {
        if (step < P.tape_k) {
            /* streaming stores: the records are read back only by the adjoint, after the
               launch (0.38 -> 0.31 ms of taping on 10M photons against cached stores) */
            uint32_t* r = P.tape + 3 * ((int64_t)step * P.tape_stride + g);
            __builtin_nontemporal_store((uint32_t)line | ((uint32_t)ev << 30), &r[0]);
            __builtin_nontemporal_store(__float_as_uint(wl), &r[1]);
            __builtin_nontemporal_store(__float_as_uint(ell), &r[2]);
        }
    }
    __device__ inline __attribute__((always_inline)) void finish(const SimParams& P, int64_t g, int steps)
    {
        P.tape_n[g] = (uint8_t)(steps < 255 ? steps : 255);
    }
};

/* record j of launch-local photon g */
OA_D const uint32_t* tape_record(const SimParams& P, int64_t g, int j)
{
    return P.tape + 3 * ((int64_t)j * P.tape_stride + g);
}

template <bool OOL = true>
OA_D void fill(const SimParams& P, State& s, uint32_t boundary, float wl, float cosTheta)
{
    const int line = (int)boundary * 4;
    const int m1_line = cosTheta > 0.f ? line + 3 : line + 0;
    const int m2_line = cosTheta > 0.f ? line + 0 : line + 3;
    const int su_line = cosTheta > 0.f ? line + 2 : line + 1;
    s.m1_line = m1_line;
    s.material1 = bnd_lookup<OOL>(P, wl, m1_line, 0);
    s.group_velocity = bnd_lookup<OOL>(P, wl, m1_line, 1).x;
    s.n2 = bnd_lookup<OOL>(P, wl, m2_line, 0).x;
    s.ems = P.optical[4 * su_line + 1];
    s.surface = (s.ems != EMS_NOSURFACE) ? bnd_lookup<OOL>(P, wl, su_line, 0) : make_float4(0.f, 0.f, 0.f, 0.f);
}

template <bool OOL = true>
OA_D float icdf_tex(const SimParams& P, float x, int layer)
{
    const int n = P.n_icdf;
    float xb = x * (float)n - 0.5f;
    float fl = floorf(xb);
    float a = xb - fl;
    int i0, i1;
    wrap_pair<OOL>((int)fl, n, i0, i1);
    int L = layer < P.n_icdf_layer ? layer : 0;
    const float* row = P.icdf + (int64_t)L * n;
    float v0 = row[i0], v1 = row[i1];
    return v0 + a * (v1 - v0);
}

template <bool OOL = true>
OA_D float scint_wavelength(const SimParams& P, float u0)
{
    if (P.icdf_hd == 0) return icdf_tex<OOL>(P, u0, 0);
    if (P.icdf_hd == 10) {
        bool lhs = u0 < 0.1f, rhs = u0 > 0.9f;
        float x = lhs ? (u0 * 10.f) : (rhs ? (u0 - 0.9f) * 10.f : u0);
        return icdf_tex<OOL>(P, x, lhs ? 1 : (rhs ? 2 : 0));
    }
    if (P.icdf_hd == 20) {
        bool lhs = u0 < 0.05f, rhs = u0 > 0.95f;
        float x = lhs ? (u0 * 20.f) : (rhs ? (u0 - 0.95f) * 20.f : u0);
        return icdf_tex<OOL>(P, x, lhs ? 1 : (rhs ? 2 : 0));
    }
    return 0.f;
}

/* ---------------------------------------------------------- helpers */

OA_D void rotateUz(v3& d, v3 u)   /* sysrap/smath.h:95-115 */
{
    float up = u.x * u.x + u.y * u.y;
    if (up > 0.f) {
        up = sqrtf(up);
        float px = d.x, py = d.y, pz = d.z;
        d.x = (u.x * u.z * px - u.y * py) / up + u.x * pz;
        d.y = (u.y * u.z * px + u.x * py) / up + u.y * pz;
        d.z = -up * px + u.z * pz;
    } else if (u.z < 0.f) {
        d.x = -d.x;
        d.z = -d.z;
    }
}

OA_D v3 uniform_sphere(float u0, float u1)
{
    float phi = u0 * 2.f * OA_PI_F;
    float cosTheta = 2.f * u1 - 1.f;
    float sinTheta = sqrtf(1.f - cosTheta * cosTheta);
    float sp, cp;
    oa_sincosf(phi, &sp, &cp);
    return mk(cp * sinTheta, sp * sinTheta, cosTheta);
}

template <class R>
OA_D void marsaglia(v3& dir, R& rng)
{
    float u, v, b, a;
    int trials = 0;
    do {
        float u0 = uni(rng);
        float u1 = uni(rng);
        u = 2.f * u0 - 1.f;
        v = 2.f * u1 - 1.f;
        b = u * u + v * v;
    } while (b > 1.f && ++trials < OA_MAX_TRIALS);
    if (b > 1.f) { u = 0.f; v = 0.f; b = 0.f; }   /* bound reached: (0, 0, -1) */
    a = 2.f * sqrtf(1.f - b);
    dir = mk(a * u, a * v, 2.f * b - 1.f);
}

template <class R>
OA_D void lambertian(v3& dir, v3 normal, float orient, R& rng)
{
    float ndotv, u;
    int count = 0;
    do {
        count++;
        marsaglia(dir, rng);
        ndotv = dot(dir, normal) * orient;
        if (ndotv < 0.f) { dir = -1.f * dir; ndotv = -1.f * ndotv; }
        u = uni(rng);
    } while (!(u < ndotv) && (count < 1024));
}

template <class R>
OA_D void rayleigh(Photon& p, R& rng)
{
    v3 direction, polarization;
    bool looping = true;
    int trials = 0;
    do {
        float u0 = uni(rng), u1 = uni(rng), u2 = uni(rng), u3 = uni(rng), u4 = uni(rng);
        float cosTheta = u0;
        float sinTheta = sqrtf(1.0f - u0 * u0);
        if (u1 < 0.5f) cosTheta = -cosTheta;
        float sinPhi, cosPhi;
        oa_sincosf(2.f * OA_PI_F * u2, &sinPhi, &cosPhi);
        direction = mk(sinTheta * cosPhi, sinTheta * sinPhi, cosTheta);
        rotateUz(direction, p.mom);
        float constant = -dot(direction, p.pol);
        polarization = mk(p.pol.x + constant * direction.x, p.pol.y + constant * direction.y,
                          p.pol.z + constant * direction.z);
        if (dot(polarization, polarization) == 0.f) {
            oa_sincosf(2.f * OA_PI_F * u3, &sinPhi, &cosPhi);
            polarization = mk(cosPhi, sinPhi, 0.f);
            rotateUz(polarization, direction);
        } else {
            if (u3 < 0.5f) polarization = -polarization;
        }
        polarization = normalize(polarization);
        float doCosTheta = dot(polarization, p.pol);
        looping = doCosTheta * doCosTheta < u4;
    } while (looping && ++trials < OA_MAX_TRIALS);
    p.mom = direction;
    p.pol = polarization;
}

/* ----------------------------------------------------------- physics */

template <class R, class Tape>
OA_D int to_boundary(const SimParams& P, uint32_t& flag, R& rng, Photon& p, const State& s, float dist, Tape& tape)
{
    const float absorption_length = s.material1.y;
    const float scattering_length = s.material1.z;
    const float reemission_prob = s.material1.w;
    const float gv = s.group_velocity;
    const float u_scattering = uni(rng);
    const float u_absorption = uni(rng);
    const float scattering_distance = -scattering_length * oa_logf(u_scattering);
    const float absorption_distance = -absorption_length * oa_logf(u_absorption);
    /* the step's one segment: its length and the process that ends it (the
       branches below, in the same order) */
    {
        const bool absorbs = absorption_distance <= scattering_distance && absorption_distance <= dist;
        const bool scatters = !(absorption_distance <= scattering_distance) && scattering_distance <= dist;
        tape.seg(P, s.m1_line, p.wavelength, absorption_length, scattering_length,
                 absorbs ? absorption_distance : scatters ? scattering_distance : dist,
                 absorbs ? SEG_ABSORB : scatters ? SEG_SCATTER : SEG_SAIL);
    }
    if (absorption_distance <= scattering_distance) {
        if (absorption_distance <= dist) {
            p.time += absorption_distance / gv;
            p.pos = p.pos + absorption_distance * p.mom;
            float u_reemit = reemission_prob == 0.f ? 2.f : uni(rng);
            if (u_reemit < reemission_prob) {
                float u_wl = uni(rng), u_mph = uni(rng), u_mct = uni(rng), u_pph = uni(rng), u_pct = uni(rng);
                p.wavelength = scint_wavelength<wrap_ool<R>::value>(P, u_wl);
                p.mom = uniform_sphere(u_mph, u_mct);
                p.pol = normalize(cross(uniform_sphere(u_pph, u_pct), p.mom));
                flag = F_BULK_REEMIT;
                return CMD_CONTINUE;
            }
            flag = F_BULK_ABSORB;
            return CMD_BREAK;
        }
    } else if (scattering_distance <= dist) {
        p.time += scattering_distance / gv;
        p.pos = p.pos + scattering_distance * p.mom;
        rayleigh(p, rng);
        flag = F_BULK_SCATTER;
        return CMD_CONTINUE;
    }
    p.pos = p.pos + dist * p.mom;
    p.time += dist / gv;
    return CMD_BOUNDARY;
}

template <class R>
OA_D int at_boundary(uint32_t& flag, R& rng, Photon& p, float n1, float n2, v3 normal)
{
    const float eta = n1 / n2;
    const float _c1 = -dot(p.mom, normal);
    const v3 on = _c1 < 0.f ? -normal : normal;
    const v3 trans = cross(p.mom, on);
    const float trans_length = length(trans);
    const bool normal_incidence = trans_length < 1e-6f;
    const v3 A_trans = normal_incidence ? p.pol : trans * (1.0f / trans_length);
    const float E1_perp = dot(p.pol, A_trans);
    const float c1 = fabsf(_c1);
    const float c2c2 = 1.f - eta * eta * (1.f - c1 * c1);
    const bool tir = c2c2 < 0.f;
    const float EdotN = dot(p.pol, on);
    const float c2 = tir ? 0.f : sqrtf(c2c2);
    const float n1c1 = n1 * c1, n2c2 = n2 * c2, n2c1 = n2 * c1, n1c2 = n1 * c2;
    float E1x, E1y;
    if (normal_incidence) { E1x = 0.f; E1y = 1.f; }
    else { E1x = E1_perp; E1y = length(p.pol - E1_perp * A_trans); }
    const float E2_tx = 2.f * n1c1 * E1x / (n1c1 + n2c2);
    const float E2_ty = 2.f * n1c1 * E1y / (n2c1 + n1c2);
    const float E2_rx = E2_tx - E1x;
    const float E2_ry = (n2 * E2_ty / n1) - E1y;
    const float invR = 1.0f / sqrtf(E2_rx * E2_rx + E2_ry * E2_ry);
    const float invT = 1.0f / sqrtf(E2_tx * E2_tx + E2_ty * E2_ty);
    const float TransCoeff = (tir || n1c1 == 0.f) ? 0.f : n2c2 * (E2_tx * E2_tx + E2_ty * E2_ty) / n1c1;
    const float u_reflect = uni(rng);
    const bool reflect = u_reflect > TransCoeff;
    p.mom = reflect ? p.mom + (2.0f * c1) * on : eta * p.mom + (eta * c1 - c2) * on;
    const v3 A_paral = normalize(cross(p.mom, A_trans));
    if (normal_incidence) {
        p.pol = reflect ? p.pol * (n2 > n1 ? -1.f : 1.f) : p.pol;
    } else if (reflect) {
        p.pol = tir ? -p.pol + (2.f * EdotN) * on : (E2_rx * invR) * A_trans + (E2_ry * invR) * A_paral;
    } else {
        p.pol = (E2_tx * invT) * A_trans + (E2_ty * invT) * A_paral;
    }
    flag = reflect ? F_BOUNDARY_REFLECT : F_BOUNDARY_TRANSMIT;
    return CMD_CONTINUE;
}

template <class R>
OA_D int at_surface(uint32_t& flag, R& rng, Photon& p, const State& s, v3 normal)
{
    const float detect = s.surface.x, absorb = s.surface.y, diffuse = s.surface.w;
    const float u_surface = uni(rng);
    const int action = u_surface < absorb + detect ? CMD_BREAK : CMD_CONTINUE;
    if (action == CMD_BREAK) {
        flag = u_surface < absorb ? F_SURFACE_ABSORB : F_SURFACE_DETECT;
    } else if (u_surface < absorb + detect + diffuse) {
        flag = F_SURFACE_DREFLECT;
        const v3 old_mom = p.mom;
        const float orient = dot(old_mom, normal) > 0.f ? -1.f : 1.f;
        lambertian(p.mom, normal, orient, rng);
        const v3 facet = normalize(p.mom - old_mom);
        const float EdotN = dot(p.pol, facet);
        p.pol = -1.f * p.pol + (2.f * EdotN) * facet;
    } else {
        flag = F_SURFACE_SREFLECT;
        const float PdotN = dot(p.mom, normal);
        p.mom = p.mom - (2.f * PdotN) * normal;
        const float EdotN = dot(p.pol, normal);
        p.pol = -1.f * p.pol + (2.f * EdotN) * normal;
    }
    return action;
}

template <class R, class Tape>
OA_D int propagate(const SimParams& P, R& rng, Photon& p, const Prd& prd, Tape& tape)
{
    const uint32_t boundary = prd.boundary;
    const uint32_t identity = prd.iindex_identity & 0xffffu;
    const uint32_t iindex = prd.iindex_identity >> 16;
    const v3 normal = prd.normal;
    const float cosTheta = dot(p.mom, normal);
    /* sphoton::set_prd (sysrap/sphoton.h:499-505) */
    p.obf = (p.obf & 0x8000ffffu) | ((boundary & 0x7fffu) << 16);
    p.identity = (p.identity & 0xff000000u) | (identity & 0x00ffffffu);
    p.obf = (p.obf & 0x7fffffffu) | ((cosTheta < 0.f ? 0x1u : 0x0u) << 31);
    p.hitcount_iindex = 0x00010000u | (0x0000ffffu & iindex);
    State s;
    fill<wrap_ool<R>::value>(P, s, boundary, p.wavelength, cosTheta);
    uint32_t flag = 0u;
    int command = to_boundary(P, flag, rng, p, s, prd.t, tape);
    if (command == CMD_BOUNDARY) {
        if (s.ems == EMS_NOSURFACE) command = at_boundary(flag, rng, p, s.material1.x, s.n2, normal);
        else if (s.ems == EMS_SURFACE) command = at_surface(flag, rng, p, s, normal);
        else if (prd.lposcost < 0.f) command = at_surface(flag, rng, p, s, normal);
        else if (s.ems == EMS_ZPLUS_SENSOR_A) { (void)uni(rng); flag = F_SURFACE_DETECT; command = CMD_BREAK; }
    }
    set_flag(p, flag);
    return command;
}

/* -------------------------------------------------------- generation */

enum : uint32_t { T_DISC = 1, T_LINE = 2, T_POINT = 3, T_CIRCLE = 4, T_RECTANGLE = 5, T_SPHERE_MARSAGLIA = 6, T_SPHERE = 7,
                  T_POINT_ISO = 100 };

OA_D void zero_flags_set(Photon& p, uint32_t flag)
{
    p.obf = 0u; p.identity = 0u; p.index = 0u; p.flagmask = 0u; p.hitcount_iindex = 0u;
    set_flag(p, flag);
}

template <class R>
OA_D void gen_torch(Photon& p, R& rng, const float4* gq, uint64_t photon_id)
{
    const float4 q0 = gq[0], q1 = gq[1], q2 = gq[2], q3 = gq[3], q4 = gq[4], q5 = gq[5];
    const uint32_t numphoton = __float_as_uint(q0.w);
    const uint32_t type = __float_as_uint(q5.w);
    const v3 gpos = mk(q1.x, q1.y, q1.z);
    const float gtime = q1.w;
    const v3 gmom = mk(q2.x, q2.y, q2.z);
    const float gweight = q2.w;
    const float gwl = q3.w;
    const float zx = q4.x, zy = q4.y, ax = q4.z, ay = q4.w;
    const float radius = q5.x, distance = q5.y;
    p.pos = mk(0.f, 0.f, 0.f); p.time = 0.f; p.mom = mk(0.f, 0.f, 0.f); p.pol = mk(0.f, 0.f, 0.f);
    p.wavelength = 0.f;
    if (type == T_SPHERE || type == T_POINT_ISO) {
        p.wavelength = gwl; p.time = gtime;
        float u_zenith = zx + uni(rng) * (zy - zx);
        float u_azimuth = ax + uni(rng) * (ay - ax);
        float phi = 2.f * OA_PI_F * u_azimuth;
        float sinPhi, cosPhi; oa_sincosf(phi, &sinPhi, &cosPhi);
        float cosTheta = 1.f - 2.0f * u_zenith;
        float sinTheta = sqrtf(1.0f - cosTheta * cosTheta);
        float flip = copysignf(1.f, radius);
        p.mom = mk(flip * sinTheta * cosPhi, flip * sinTheta * sinPhi, flip * cosTheta);
        float rad = fabsf(radius);
        p.pos = mk(sinTheta * cosPhi * rad, sinTheta * sinPhi * rad, cosTheta * rad);
        float phase = 2.f * OA_PI_F * distance;
        float sph, cph; oa_sincosf(phase, &sph, &cph);
        p.pol = mk(cph, sph, 0.f);
        rotateUz(p.pol, p.mom);
        if (type == T_POINT_ISO) {
            p.pos = p.pos + gpos;
            if (gweight > 0.f) p.wavelength = gwl + uni(rng) * gweight;
        }
    } else if (type == T_DISC) {
        p.wavelength = gwl; p.time = gtime; p.mom = gmom;
        float u_zenith = zx + uni(rng) * (zy - zx);
        float u_azimuth = ax + uni(rng) * (ay - ax);
        float r = radius * u_zenith;
        float phi = 2.f * OA_PI_F * u_azimuth;
        float sinPhi, cosPhi; oa_sincosf(phi, &sinPhi, &cosPhi);
        p.pos = mk(r * cosPhi, r * sinPhi, 0.f);
        rotateUz(p.pos, p.mom);
        p.pos = p.pos + gpos;
        p.pol = mk(sinPhi, -cosPhi, 0.f);
        rotateUz(p.pol, p.mom);
    } else if (type == T_SPHERE_MARSAGLIA) {
        p.wavelength = gwl; p.time = gtime;
        float u, v, b, a;
        int trials = 0;
        do {
            float u0z = zx + uni(rng) * (zy - zx);
            float u1a = ax + uni(rng) * (ay - ax);
            u = 2.f * u0z - 1.f; v = 2.f * u1a - 1.f; b = u * u + v * v;
        } while (b > 1.f && ++trials < OA_MAX_TRIALS);
        if (b > 1.f) { u = 0.f; v = 0.f; b = 0.f; }
        a = 2.f * sqrtf(1.f - b);
        float rad = fabsf(radius);
        float flip = copysignf(1.f, radius);
        p.mom = mk(flip * a * u, flip * a * v, flip * (2.f * b - 1.f));
        p.pos = mk(a * u * rad, a * v * rad, (2.f * b - 1.f) * rad);
        float phase = 2.f * OA_PI_F * distance;
        float sph, cph; oa_sincosf(phase, &sph, &cph);
        p.pol = mk(cph, sph, 0.f);
        rotateUz(p.pol, p.mom);
    } else if (type == T_LINE) {
        p.wavelength = gwl; p.time = gtime; p.mom = gmom;
        float frac = (float)photon_id / (float)numphoton;
        float sfrac = 2.f * (frac - 0.5f);
        float r = radius * sfrac;
        p.pos = mk(r, 0.f, 0.f);
        rotateUz(p.pos, p.mom);
        p.pos = p.pos + gpos;
        p.pol = mk(0.f, -1.f, 0.f);
        rotateUz(p.pol, p.mom);
    } else if (type == T_POINT) {
        p.wavelength = gwl; p.time = gtime; p.mom = gmom;
        p.pos = mk(0.f, 0.f, 0.f);
        rotateUz(p.pos, p.mom);
        p.pos = p.pos + gpos;
        p.pol = mk(0.f, -1.f, 0.f);
        rotateUz(p.pol, p.mom);
    } else if (type == T_CIRCLE) {
        p.wavelength = gwl; p.time = gtime;
        float f = (float)photon_id / (float)numphoton;
        float frac = ax * (1.f - f) + ay * (f);
        float phi = 2.f * OA_PI_F * frac;
        float sinPhi, cosPhi; oa_sincosf(phi, &sinPhi, &cosPhi);
        float r = radius < 0.f ? -radius : radius;
        p.mom = mk(radius < 0.f ? -cosPhi : cosPhi, 0.f, radius < 0.f ? -sinPhi : sinPhi);
        p.pos = mk(r * cosPhi, 0.f, r * sinPhi);
        p.pos = p.pos + gpos;
        p.pol = mk(0.f, -1.f, 0.f);
        rotateUz(p.pol, p.mom);
    } else if (type == T_RECTANGLE) {
        /* storch.h:463-516: photons split over the four sides of the rectangle
           x in [azimuth.x, azimuth.y], z in [zenith.x, zenith.y] of the y = 0 plane,
           each pointing inwards; left, right, bottom, top. Fewer than 4 photons
           (side size 0, a division by zero in the reference) all go to side 0. */
        p.wavelength = gwl; p.time = gtime;
        const int side_size = (int)(numphoton / 4u);
        const uint64_t ss = side_size > 0 ? (uint64_t)side_size : 1ull;
        const int side = (int)(photon_id / ss);
        const int side_index = (int)(photon_id - (uint64_t)side * ss);
        const float frac = (float)side_index / (float)(side_size > 0 ? side_size : 1);
        if (side == 0 || side == 1) {
            p.pos = mk(side == 0 ? ax : ay, 0.f, (1.f - frac) * zx + frac * zy);
            p.mom = mk(side == 0 ? 1.f : -1.f, 0.f, 0.f);
        } else if (side == 2 || side == 3) {
            p.pos = mk((1.f - frac) * ax + frac * ay, 0.f, side == 2 ? zx : zy);
            p.mom = mk(0.f, 0.f, side == 2 ? 1.f : -1.f);
        }
        p.pos = p.pos + gpos;
        p.pol = mk(0.f, -1.f, 0.f);
        rotateUz(p.pol, p.mom);
    }
    zero_flags_set(p, F_TORCH);
}

template <class R>
OA_D void gen_scint(const SimParams& P, Photon& p, R& rng, const float4* gq)
{
    float u0 = uni(rng), u1 = uni(rng), u2 = uni(rng), u3 = uni(rng);
    float cost = 1.f - 2.f * u0;
    float sint = sqrtf((1.f - cost) * (1.f + cost));
    float phi = 2.f * OA_PI_F * u1;
    float sinp, cosp; oa_sincosf(phi, &sinp, &cosp);
    p.mom = mk(sint * cosp, sint * sinp, cost);
    p.pol = mk(cost * cosp, cost * sinp, -sint);
    phi = 2.f * OA_PI_F * u2;
    oa_sincosf(phi, &sinp, &cosp);
    p.pol = normalize(cosp * p.pol + sinp * cross(p.mom, p.pol));
    p.wavelength = scint_wavelength<wrap_ool<R>::value>(P, u3);
    const float4 q1 = gq[1], q2 = gq[2], q3 = gq[3], q5 = gq[5];
    const float charge = q3.y, meanVelocity = q3.w, step_length = q2.w, scint_time = q5.x;
    float fraction = charge == 0.f ? 1.f : uni(rng);
    p.pos = mk(q1.x, q1.y, q1.z) + fraction * mk(q2.x, q2.y, q2.z);
    float u4 = uni(rng);
    float deltaTime = fraction * step_length / meanVelocity - scint_time * oa_logf(u4);
    p.time = q1.w + deltaTime;
    zero_flags_set(p, F_SCINTILLATION);
}

/* qcerenkov::generate (qudarap/qcerenkov.h): wavelength by reciprocal sampling
   flat in energy with rejection on sin^2(theta) against the medium's RINDEX
   from the bnd table (line = genstep matline), cone about the normalized step,
   emission point along the step by rejection on the linear N1 -> N2 yield */
template <class R>
OA_D void gen_cerenkov(const SimParams& P, Photon& p, R& rng, const float4* gq)
{
    const float4 q0 = gq[0], q1 = gq[1], q2 = gq[2], q3 = gq[3], q4 = gq[4], q5 = gq[5];
    const int matline = (int)__float_as_uint(q0.z);
    const v3 dpos = mk(q2.x, q2.y, q2.z);
    const float step_length = q2.w, preVelocity = q3.w;
    const float BetaInverse = q4.x, Wmin = q4.y, Wmax = q4.z;
    const float maxSin2 = q5.x, N1 = q5.y, N2 = q5.z, postVelocity = q5.w;
    const v3 p0 = normalize(dpos);
    float wavelength, cosTheta, sin2Theta, u_maxSin2;
    unsigned count = 0;
    do {
        const float u0 = uni(rng);
        const float w = Wmin + u0 * (Wmax - Wmin);
        wavelength = Wmin * Wmax / w;
        const float4 props = bnd_lookup<wrap_ool<R>::value>(P, wavelength, matline, 0);
        const float sampledRI = props.x;
        cosTheta = BetaInverse / sampledRI;
        sin2Theta = fmaxf(0.f, (1.f - cosTheta) * (1.f + cosTheta));
        const float u1 = uni(rng);
        u_maxSin2 = u1 * maxSin2;
        count += 1;
    } while (u_maxSin2 > sin2Theta && count < 100);
    const float sinTheta = sqrtf(sin2Theta);
    const float u0 = uni(rng);
    const float phi = 2.f * OA_PI_F * u0;
    float sinPhi, cosPhi;
    oa_sincosf(phi, &sinPhi, &cosPhi);
    p.mom = mk(sinTheta * cosPhi, sinTheta * sinPhi, cosTheta);
    rotateUz(p.mom, p0);
    p.pol = mk(cosTheta * cosPhi, cosTheta * sinPhi, -sinTheta);
    rotateUz(p.pol, p0);
    p.wavelength = wavelength;
    float fraction, delta, NumberOfPhotons, N, u;
    const float MeanNumberOfPhotonsMax = fmaxf(N1, N2);
    const float DeltaN = (N1 - N2);
    int trials = 0;
    do {
        fraction = uni(rng);
        delta = fraction * step_length;
        NumberOfPhotons = N1 - fraction * DeltaN;
        u = uni(rng);
        N = u * MeanNumberOfPhotonsMax;
    } while (N > NumberOfPhotons && ++trials < OA_MAX_TRIALS);
    const float midVelocity = preVelocity + fraction * (postVelocity - preVelocity) * 0.5f;
    p.time = q1.w + delta / midVelocity;
    p.pos = mk(q1.x, q1.y, q1.z) + fraction * dpos;
    zero_flags_set(p, F_CERENKOV);
}

/* ------------------------------------------------------------- trace */

/* slab test; on a hit tn is the entry distance clamped to tmin */
/* Ray data for slab tests: t = fma(plane, inv, -oi) with inv ~ 1/d (the hardware
   reciprocal, 1 ulp) and oi = o*inv. A direction component below 1e-20 in
   magnitude is replaced by +-1e-20 so inv and oi stay finite and the slab of a
   parallel ray gives huge t of the right sign. The reciprocal and the FMA form
   round differently from (plane - o)/d by a few ulp of o; box tests only prune
   (every box is padded by >= 1e-2 mm around what it holds), so the hits found
   do not depend on it. */
OA_D void slab_ray(v3 o, v3 d, v3& inv, v3& oi)
{
    const float e = 1e-20f;
    inv = mk(__builtin_amdgcn_rcpf(fabsf(d.x) < e ? copysignf(e, d.x) : d.x),
             __builtin_amdgcn_rcpf(fabsf(d.y) < e ? copysignf(e, d.y) : d.y),
             __builtin_amdgcn_rcpf(fabsf(d.z) < e ? copysignf(e, d.z) : d.z));
    oi = mk(o.x * inv.x, o.y * inv.y, o.z * inv.z);
}

/* entry distance into one child box, or +inf when the ray misses it in [tmin, tmax] */
OA_D float box_enter(float lx, float ly, float lz, float hx, float hy, float hz, v3 inv, v3 oi, float tmin,
                     float tmax)
{
    const float x0 = fmaf(lx, inv.x, -oi.x), x1 = fmaf(hx, inv.x, -oi.x);
    const float y0 = fmaf(ly, inv.y, -oi.y), y1 = fmaf(hy, inv.y, -oi.y);
    const float z0 = fmaf(lz, inv.z, -oi.z), z1 = fmaf(hz, inv.z, -oi.z);
    const float tn = fmaxf(fmaxf(fminf(x0, x1), fminf(y0, y1)), fmaxf(fminf(z0, z1), tmin));
    const float tf = fminf(fminf(fmaxf(x0, x1), fmaxf(y0, y1)), fminf(fmaxf(z0, z1), tmax));
    return tn <= tf ? tn : INFINITY;
}

OA_D void cswap(float& ta, uint32_t& ca, float& tb, uint32_t& cb)
{
    const bool s = tb < ta;
    const float t = s ? tb : ta;
    const uint32_t c = s ? cb : ca;
    tb = s ? ta : tb; cb = s ? ca : cb;
    ta = t; ca = c;
}

/* Visit one 4-wide node: returns the nearest hit child link (or EMPTY) and
   pushes the other hit children farthest first, so they pop nearest first.
   Leaf links go on the stack like node links, so both levels are traversed
   nearest-first and a shrinking tmax culls what is still pending. Equal entry
   distances keep the node's child order (leaves first). */
template <int STRIDE, bool SINGLE = false>
OA_D uint32_t visit(const BvhNode* nodes, uint32_t idx, v3 inv, v3 oi, float tmin, float tmax,
                    uint32_t* stk, int& sp)
{
    const float4* q = reinterpret_cast<const float4*>(nodes + idx);
    const float4 lx = q[0], ly = q[1], lz = q[2], hx = q[3], hy = q[4], hz = q[5];
    const uint4 ch = *reinterpret_cast<const uint4*>(q + 6);
    if constexpr (SINGLE) {
        /* a node with one child (the BLAS root of a one-prim solid): one box test.
           Empty slots hold boxes at +inf, which every test misses, so the full
           visit below would return the same link and push nothing. */
        if (ch.y == OA_BVH_EMPTY) {
            const float t = box_enter(lx.x, ly.x, lz.x, hx.x, hy.x, hz.x, inv, oi, tmin, tmax);
            return t < INFINITY ? ch.x : OA_BVH_EMPTY;
        }
    }
    float t0 = box_enter(lx.x, ly.x, lz.x, hx.x, hy.x, hz.x, inv, oi, tmin, tmax);
    float t1 = box_enter(lx.y, ly.y, lz.y, hx.y, hy.y, hz.y, inv, oi, tmin, tmax);
    float t2 = box_enter(lx.z, ly.z, lz.z, hx.z, hy.z, hz.z, inv, oi, tmin, tmax);
    float t3 = box_enter(lx.w, ly.w, lz.w, hx.w, hy.w, hz.w, inv, oi, tmin, tmax);
    uint32_t c0 = ch.x, c1 = ch.y, c2 = ch.z, c3 = ch.w;
    cswap(t0, c0, t1, c1);
    cswap(t2, c2, t3, c3);
    cswap(t0, c0, t2, c2);
    cswap(t1, c1, t3, c3);
    cswap(t1, c1, t2, c2);
    if (t3 < INFINITY) { stk[sp * STRIDE] = c3; sp++; }
    if (t2 < INFINITY) { stk[sp * STRIDE] = c2; sp++; }
    if (t1 < INFINITY) { stk[sp * STRIDE] = c1; sp++; }
    return t0 < INFINITY ? c0 : OA_BVH_EMPTY;
}

OA_D void inst_ray(const float4* m, v3 o, v3 d, v3& oo, v3& od)
{
    const float4 r0 = m[0], r1 = m[1], r2 = m[2], r3 = m[3];
    oo = mk(r0.x * o.x + r1.x * o.y + r2.x * o.z + r3.x,
            r0.y * o.x + r1.y * o.y + r2.y * o.z + r3.y,
            r0.z * o.x + r1.z * o.y + r2.z * o.z + r3.z);
    od = mk(r0.x * d.x + r1.x * d.y + r2.x * d.z,
            r0.y * d.x + r1.y * d.y + r2.y * d.z,
            r0.z * d.x + r1.z * d.y + r2.z * d.z);
}

/* one prim of one instance (entry k of the BLAS-ordered prim records) against the
   nearest hit so far (t, then the (instance, prim) key) */
template <bool EXT>
OA_D void consider_prim(const SimParams& P, uint32_t ii, uint32_t k, v3 oo, v3 od, float tmin, float& best_t,
                        uint64_t& best_key, uint32_t& best_inst, uint32_t& best_prim, Isect& best_is)
{
    const PrimRec pr = P.blas_prims[k];
    const uint32_t pi = pr.index;
    Isect is;
    if (!intersect_prim<EXT>(is, P.node + 4u * pr.node_offset, P.itra, P.plan, tmin, oo, od)) return;
    /* a reported hit outside [tmin, tmax] is ignored, as optixReportIntersection does
       (a convex polyhedron behind the ray reports valid with t = 0) */
    if (!(is.w >= tmin)) return;
    const uint64_t key = ((uint64_t)ii << 32) | (uint64_t)pi;
    if (is.w < best_t || (is.w == best_t && key < best_key)) {
        best_t = is.w; best_key = key; best_inst = ii; best_prim = pi; best_is = is;
    }
}

/*
 * Nearest hit over instances x prims with t in (tmin, tmax); exact ties go to
 * the smaller (instance, prim) key so the answer does not depend on traversal
 * order. stk: this lane's LDS stack, entry k at stk[k*STRIDE].
 *
 * "while-while" traversal (Aila and Laine, "Understanding the efficiency of
 * ray traversal on GPUs", HPG 2009): lanes walk TLAS nodes until each holds
 * an instance leaf or is finished, then all pending leaves (instance ray
 * transform, BLAS walk, CSG prim tests) run together, so the expensive CSG
 * code executes with most lanes of the wave active instead of once per lane.
 */
template <int STRIDE, bool EXT = false>
OA_D void trace(const SimParams& P, uint32_t* stk, v3 o, v3 d, float tmin, float tmax, Prd& prd)
{
    OA_COUNT(OA_ST_TRACE);
    OA_WAVE_COUNT(OA_ST_W_TRACE, OA_ST_L_TRACE);
    OA_CYC_MARK(c_trace);
    v3 inv, oi;
    slab_ray(o, d, inv, oi);
    float best_t = tmax;
    uint64_t best_key = ~0ull;
    uint32_t best_inst = 0xffffffffu, best_prim = 0u;
    Isect best_is; best_is.x = best_is.y = best_is.z = best_is.w = 0.f;
    int sp = 0;
    uint32_t cur = P.tlas_empty ? OA_BVH_EMPTY : 0u;    /* TLAS root node */
    /* the first leaf phase tests the front instances (world-sized boxes kept out
       of the TLAS) on every lane together; the walk then starts with their
       nearest hit as tmax */
    uint32_t leaf = P.front_leaf;
    /* lower bound of the node and BLAS box tests of TLAS instances (set after the front phase) */
    float tlo = tmin;
    bool done = false;
#ifdef OA_STATS
    uint32_t n1 = 0, n2 = 0;
#endif
    while (!done) {
        /* phase 1: walk nodes until this lane has a leaf or nothing left */
        OA_CYC_MARK(c_p1);
        /* Once fewer than regroup_pct percent of the lanes that entered this walk
           still walk, the whole wave leaves it: the lanes holding a leaf run the
           leaf phase and the others resume from their own node and stack after it.
           Each lane still visits the same nodes and leaves in the same order, so
           only the wave's interleaving changes, never a result. 0 waits for all. */
        const int walk_floor = P.regroup_pct > 0 ? __popcll(__ballot(1)) * P.regroup_pct : 0;
        while (!done && leaf == OA_BVH_EMPTY) {
            if (__popcll(__ballot(1)) * 100 < walk_floor) break;
            OA_WAVE_COUNT(OA_ST_W_P1, OA_ST_L_P1);
#ifdef OA_STATS
            n1++;
#endif
            /* one iteration: pop when the lane holds nothing, visit a node, and
               take a leaf (popped or the visit's nearest child) at once */
            if (cur == OA_BVH_EMPTY) {
                if (sp == 0) { done = true; break; }
                sp--;
                cur = stk[sp * STRIDE];
            }
            if (!(cur & OA_BVH_LEAF_BIT)) {
                OA_COUNT(OA_ST_TLAS);
#ifdef OA_STATS
                const int sp0 = sp;
#endif
                cur = visit<STRIDE>(P.tlas, cur, inv, oi, tlo, best_t, stk, sp);
#ifdef OA_STATS
                if (cur == OA_BVH_EMPTY && sp == sp0) OA_COUNT(OA_ST_TLAS_DEAD);
#endif
            }
            if (cur != OA_BVH_EMPTY && (cur & OA_BVH_LEAF_BIT)) {
                leaf = cur;
                cur = OA_BVH_EMPTY;
            }
        }
        OA_CYC_ADD(OA_ST_C_P1, c_p1);
        if (leaf == OA_BVH_EMPTY) continue;
        /* phase 2: the pending instance leaf */
        OA_CYC_MARK(c_p2);
        /* the leaf phase is chains of dependent loads (instance record, BLAS node, prim
           record, CSG nodes): waves in it issue first, so their next load starts sooner
           (0.6% on the 64 events against no priority; raising the walk instead lost 1%) */
        __builtin_amdgcn_s_setprio(2);
        const uint32_t first = leaf & 0x00ffffffu, count = (leaf >> 24) & 0x7fu;
        leaf = OA_BVH_EMPTY;
        for (uint32_t k = 0; k < count; ++k) {
            const uint32_t ii = P.tlas_items[first + k];
            OA_COUNT(OA_ST_INST);
            const float4* m = reinterpret_cast<const float4*>(P.inst + ii);
            v3 oo, od;
            inst_ray(m, o, d, oo, od);
            v3 inv2, oi2;
            slab_ray(oo, od, inv2, oi2);
            const int sp_base = sp;
            uint32_t bc = __float_as_uint(m[2].w);     /* BLAS root of this solid (front instances: one leaf of all its prims) */
            for (;;) {
                OA_WAVE_COUNT(OA_ST_W_P2, OA_ST_L_P2);
#ifdef OA_STATS
                n2++;
#endif
                /* as in the node walk: pop, visit, and test a leaf's prims in one iteration */
                if (bc == OA_BVH_EMPTY) {
                    if (sp == sp_base) break;
                    sp--;
                    bc = stk[sp * STRIDE];
                }
                if (!(bc & OA_BVH_LEAF_BIT)) {
                    OA_COUNT(OA_ST_BLAS);
#ifdef OA_STATS
                    const int sp0 = sp;
#endif
                    bc = visit<STRIDE, true>(P.blas, bc, inv2, oi2, tlo, best_t, stk, sp);
#ifdef OA_STATS
                    if (bc == OA_BVH_EMPTY && sp == sp0) OA_COUNT(OA_ST_BLAS_DEAD);
#endif
                    if (bc == OA_BVH_EMPTY || !(bc & OA_BVH_LEAF_BIT)) continue;
                }
                const uint32_t bf = bc & 0x00ffffffu, bn = (bc >> 24) & 0x7fu;
                OA_CYC_MARK(c_prim);
                for (uint32_t j = 0; j < bn; ++j) {
                    OA_COUNT(OA_ST_PRIM);
                    OA_WAVE_COUNT(OA_ST_W_PRIM, OA_ST_L_PRIM);
                    consider_prim<EXT>(P, ii, bf + j, oo, od, tmin, best_t, best_key, best_inst, best_prim, best_is);
                }
                OA_CYC_ADD(OA_ST_C_PRIM, c_prim);
                bc = OA_BVH_EMPTY;
            }
        }
        OA_CYC_ADD(OA_ST_C_P2, c_p2);
        __builtin_amdgcn_s_setprio(0);
        /* after the front phase (the TLAS walk not begun: cur is still the root): a lane
           whose segment from the origin to the nearest front hit lies in the free ball,
           which no TLAS instance meets, has nothing to find in the TLAS; one whose origin
           lies in it can meet an instance only after leaving it, so its box tests start
           there (boxes only prune: the prim tests keep tmin) */
        if (cur == 0u) {
            const float4 fb = *P.free_ball;
            if (fb.w > 0.f) {
                const v3 c = mk(fb.x, fb.y, fb.z);
                const v3 a = o - c, b = o + best_t * d - c;
                const float qa = dot(a, a) - fb.w;
                if (qa < 0.f) {
                    if (dot(b, b) < fb.w) {
                        cur = OA_BVH_EMPTY;
                    } else {
                        const float dd = dot(d, d), bb = dot(a, d);
                        const float t_exit = (sqrtf(bb * bb - dd * qa) - bb) / dd;
                        tlo = fmaxf(tmin, t_exit);
                    }
                }
            }
        }
    }
    OA_CYC_ADD(OA_ST_C_TRACE, c_trace);
#ifdef OA_STATS
    atomicAdd(&oa_hist[0][n1 ? 32 - __clz(n1) : 0], 1ull);
    atomicAdd(&oa_hist[1][n2 ? 32 - __clz(n2) : 0], 1ull);
    if (n1 > OA_SLOW_ITER) {
        const unsigned long long k = atomicAdd(&oa_slow_n, 1ull);
        if (k < OA_SLOW_MAX) {
            oa_slow[3 * k + 0] = make_float4(o.x, o.y, o.z, tmin);
            oa_slow[3 * k + 1] = make_float4(d.x, d.y, d.z, tmax);
            oa_slow[3 * k + 2] = make_float4(best_t, (float)n1, (float)n2, __uint_as_float(best_inst));
        }
    }
#endif
    if (best_inst == 0xffffffffu) {
        prd.normal = mk(0.f, 0.f, 0.f); prd.t = 0.f; prd.lposcost = 0.f;
        prd.lx = 0.f; prd.ly = __uint_as_float(0x7fc00000u); prd.gprim = 0xffffu;
        prd.iindex_identity = 0xffffffffu; prd.boundary = 0xffffu;
        return;
    }
    const float4* m = reinterpret_cast<const float4*>(P.inst + best_inst);
    const float4 r0 = m[0], r1 = m[1], r2 = m[2];
    prd.normal = mk(r0.x * best_is.x + r0.y * best_is.y + r0.z * best_is.z,
                    r1.x * best_is.x + r1.y * best_is.y + r1.z * best_is.z,
                    r2.x * best_is.x + r2.y * best_is.y + r2.z * best_is.z);
    prd.t = best_t;
    v3 oo, od;
    inst_ray(m, o, d, oo, od);
    const v3 lpos = oo + best_t * od;
    prd.lposcost = lpos.z / sqrtf(dot(lpos, lpos));
    prd.lx = lpos.x;
    prd.ly = lpos.y;
    const uint32_t ident = __float_as_uint(r1.w);
    prd.iindex_identity = ((best_inst & 0xffffu) << 16) | (ident & 0xffffu);
    prd.boundary = P.prim[best_prim].boundary & 0xffffu;
    prd.gprim = P.prim[best_prim].global_prim;
}

} // namespace oa
