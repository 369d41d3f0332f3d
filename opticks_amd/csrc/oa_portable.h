/*
 * oa_portable.h : bit-reproducible FP32 transcendentals, the Philox4x32-10
 * counter-based generator and the xorwow generator, shared by the HIP kernels
 * and by the CPU checker.
 *
 * Everything here is plain C99 arithmetic (no libm calls apart from floorf and
 * sqrtf, which are correctly rounded on both targets) so that gcc and hipcc,
 * both invoked with -ffp-contract=off, produce the same bits.
 *
 * Transcendentals (written from the mathematics, methods named per function):
 *   oa_logf    Cody-Waite reduction x = 2^k m, m in [sqrt(1/2), sqrt(2)),
 *              log(m) = 2 atanh(s), s = (m-1)/(m+1), evaluated in the
 *              "f - (hfsq - s(hfsq+R))" form of Sun's fdlibm e_logf.c with
 *              R the truncated Taylor series of 2 atanh(s)/s - 2
 *              (coefficients 2/(2i+1)); ln2 split into a 12-bit head and a tail.
 *   oa_sinf / oa_cosf / oa_sincosf
 *              Cody-Waite reduction by pi/2 (three-part split, 12+12+24 bits)
 *              and Taylor polynomials on [-pi/4, pi/4] (sin to r^9, cos to r^10).
 *   oa_atan2f  octant reduction to a = min(|x|,|y|)/max(|x|,|y|) in [0,1], then
 *              atan(a) = pi/6 + atan((a sqrt3 - 1)/(a + sqrt3)) for a > tan(pi/12)
 *              so the argument t has |t| <= tan(pi/12), atan(t) = t P(t^2) with P
 *              a degree-4 least-squares fit of atan(t)/t on [0, tan(pi/12)]
 *              (fitted for this file: <= 1.02 ulp on that interval), and the
 *              quadrant reconstruction pi/2 - r, pi - r, -r. Signed zeros are
 *              not distinguished (atan2(0, -0) gives 0).
 *   Valid domains: oa_logf for positive finite normal x; sin/cos for |x| < 1e4.
 *   Accuracy is pinned against glibc in tests/test_portable_math.py.
 *
 * Random numbers:
 *   Philox4x32-10 block function from J. Salmon, M. Moraes, R. Dror, D. Shaw,
 *   "Parallel random numbers: as easy as 1, 2, 3", SC'11 (multipliers
 *   0xD2511F53/0xCD9E8D57, Weyl key increments 0x9E3779B9/0xBB67AE85).
 *   Stream convention of this project (documented, not cuRAND's):
 *     key     = (seed_lo32, seed_hi32)
 *     counter = (block_lo32, block_hi32, photon_lo32, photon_hi32)
 *   output word j of a photon's stream is word (j%4) of block j/4.
 *   The reference keys each photon's stream on the absolute photon index and
 *   separates events by a skipahead of event_idx*skipahead_event_offset words
 *   (qudarap/qrng.h:131-137); oa_rng_init reproduces that structure.
 *   Uniform mapping: u = ((x >> 8) + 1) * 2^-24, exactly representable, in
 *   (0, 1] like curand_uniform's documented range (logf(u) must be finite).
 */
#ifndef OA_PORTABLE_H
#define OA_PORTABLE_H

#include <stdint.h>

#if defined(__HIPCC__)
#define OA_FN __host__ __device__ static inline __attribute__((always_inline))
#else
#define OA_FN static inline
#endif

#if !defined(__HIPCC__)
#include <math.h>
#endif

OA_FN uint32_t oa_f2u(float f) { union { float f; uint32_t u; } c; c.f = f; return c.u; }
OA_FN float oa_u2f(uint32_t u) { union { float f; uint32_t u; } c; c.u = u; return c.f; }

/* ------------------------------------------------------------------ log */

#define OA_LN2_HI   0.693115234375f            /* 0x3f317000: ln2 to 12 bits  */
#define OA_LN2_LO   3.1946183298714459e-05f    /* 0x3805fdf4: ln2 - OA_LN2_HI */
#define OA_SQRT2_F  1.41421353816986083984f    /* 0x3fb504f3                  */

OA_FN float oa_logf(float x)
{
    uint32_t ix = oa_f2u(x);
    int k = (int)(ix >> 23) - 127;
    float m = oa_u2f((ix & 0x007fffffu) | 0x3f800000u);   /* m in [1,2) */
    if (m > OA_SQRT2_F) { m = m * 0.5f; k = k + 1; }      /* exact       */
    float f = m - 1.0f;                                   /* exact (Sterbenz) */
    float s = f / (2.0f + f);
    float z = s * s;
    float R = z * (0.666666686534881591796875f             /* 2/3  */
            + z * (0.4000000059604644775390625f            /* 2/5  */
            + z * (0.2857142984867095947265625f            /* 2/7  */
            + z * (0.2222222238779067993164062f            /* 2/9  */
            + z * (0.1818181872367858886718750f)))));      /* 2/11 */
    float hfsq = 0.5f * f * f;
    float dk = (float)k;
    return dk * OA_LN2_HI - ((hfsq - (s * (hfsq + R) + dk * OA_LN2_LO)) - f);
}

/* -------------------------------------------------------------- sin/cos */

#define OA_TWO_OVER_PI 0.636619746685028076171875f  /* 0x3f22f983 */
#define OA_PIO2_1 1.5703125f                        /* 0x3fc90000 */
#define OA_PIO2_2 4.8375129699707031e-04f           /* 0x39fda000 */
#define OA_PIO2_3 7.5497901264043321e-08f           /* 0x33a22169 */

OA_FN float oa_sin_poly(float r)
{
    float r2 = r * r;
    return r + r * r2 * (-0.16666667163372039794921875f          /* -1/3!  */
             + r2 * (8.33333376795053482055664062e-03f           /*  1/5!  */
             + r2 * (-1.98412701138295233249664307e-04f          /* -1/7!  */
             + r2 * (2.75573188446287531405687332e-06f))));      /*  1/9!  */
}

OA_FN float oa_cos_poly(float r)
{
    float r2 = r * r;
    return 1.0f - 0.5f * r2 + r2 * r2 * (4.16666679084300994873046875e-02f  /*  1/4!  */
             + r2 * (-1.38888892251998186111450195e-03f                    /* -1/6!  */
             + r2 * (2.48015876569843292236328125e-05f                     /*  1/8!  */
             + r2 * (-2.75573199814971303567290306e-07f))));               /* -1/10! */
}

OA_FN float oa_reduce_pio2(float x, int* quadrant)
{
    float kf = floorf(x * OA_TWO_OVER_PI + 0.5f);
    *quadrant = ((int)kf) & 3;
    return ((x - kf * OA_PIO2_1) - kf * OA_PIO2_2) - kf * OA_PIO2_3;
}

OA_FN void oa_sincosf(float x, float* sn, float* cs)
{
    int q;
    float r = oa_reduce_pio2(x, &q);
    float s = oa_sin_poly(r);
    float c = oa_cos_poly(r);
    float so = (q == 0) ? s : (q == 1) ? c : (q == 2) ? -s : -c;
    float co = (q == 0) ? c : (q == 1) ? -s : (q == 2) ? -c : s;
    *sn = so;
    *cs = co;
}

OA_FN float oa_sinf(float x) { float s, c; oa_sincosf(x, &s, &c); return s; }
OA_FN float oa_cosf(float x) { float s, c; oa_sincosf(x, &s, &c); return c; }

#define OA_PORT_PI_F 3.14159265358979323846f
#define OA_PI_6_F   0.52359877559829887308f
#define OA_PI_2_F   1.57079632679489661923f
#define OA_TAN_PI_12_F 0.26794919243112270647f
#define OA_SQRT3_F  1.73205080756887729353f

OA_FN float oa_atan2f(float y, float x)
{
    const float ax = fabsf(x), ay = fabsf(y);
    const float mx = fmaxf(ax, ay), mn = fminf(ax, ay);
    const float a = mx > 0.f ? mn / mx : 0.f;
    const int big = a > OA_TAN_PI_12_F;
    const float t = big ? (a * OA_SQRT3_F - 1.0f) / (a + OA_SQRT3_F) : a;
    const float z = t * t;
    float p = 0.0966006294f;                 /* 0x3dc5d68d */
    p = p * z + -0.142003193f;               /* 0xbe116949 */
    p = p * z + 0.199979603f;                /* 0x3e4cc774 */
    p = p * z + -0.333333164f;               /* 0xbeaaaaa5 */
    p = p * z + 1.0f;
    float r = t * p;
    r = big ? r + OA_PI_6_F : r;
    r = ay > ax ? OA_PI_2_F - r : r;
    r = x < 0.f ? OA_PORT_PI_F - r : r;
    return y < 0.f ? -r : r;
}

/* sphotonlite lposfphi: azimuth of a local position as a fraction of a turn
   (normalize_fphi, sysrap/scuda.h:670-673) */
OA_FN float oa_fphi(float y, float x) { return (oa_atan2f(y, x) + OA_PORT_PI_F) / (2.0f * OA_PORT_PI_F); }

/* --------------------------------------------------------------- Philox */

#define OA_PHILOX_M0 0xD2511F53u
#define OA_PHILOX_M1 0xCD9E8D57u
#define OA_PHILOX_W0 0x9E3779B9u
#define OA_PHILOX_W1 0xBB67AE85u

typedef struct { uint32_t v[4]; } oa_u32x4;

OA_FN void oa_mulhilo(uint32_t a, uint32_t b, uint32_t* hi, uint32_t* lo)
{
    uint64_t p = (uint64_t)a * (uint64_t)b;
    *hi = (uint32_t)(p >> 32);
    *lo = (uint32_t)p;
}

/* One Philox4x32-10 block: 10 rounds with 9 key bumps in between. */
OA_FN oa_u32x4 oa_philox4x32_10(oa_u32x4 ctr, uint32_t k0, uint32_t k1)
{
    uint32_t c0 = ctr.v[0], c1 = ctr.v[1], c2 = ctr.v[2], c3 = ctr.v[3];
    for (int r = 0; r < 10; ++r) {
        uint32_t hi0, lo0, hi1, lo1;
        oa_mulhilo(OA_PHILOX_M0, c0, &hi0, &lo0);
        oa_mulhilo(OA_PHILOX_M1, c2, &hi1, &lo1);
        uint32_t n0 = hi1 ^ c1 ^ k0;
        uint32_t n2 = hi0 ^ c3 ^ k1;
        c0 = n0; c1 = lo1; c2 = n2; c3 = lo0;
        k0 += OA_PHILOX_W0;
        k1 += OA_PHILOX_W1;
    }
    oa_u32x4 out;
    out.v[0] = c0; out.v[1] = c1; out.v[2] = c2; out.v[3] = c3;
    return out;
}

OA_FN float oa_u32_to_uniform(uint32_t x)
{
    return (float)((x >> 8) + 1u) * 5.9604644775390625e-08f;   /* 2^-24 */
}

/*
 * Per-photon stream. The unread words of the current block sit in b0..b3 as a
 * FIFO (next word always in b0), so no word is ever selected by a run-time
 * index: the HIP compiler keeps the whole stream in VGPRs. The next block is
 * generated as soon as the current one is used up.
 */
typedef struct {
    uint64_t key;             /* seed: key words (lo, hi)                  */
    uint64_t photon;          /* photon index: counter words 2,3 (lo, hi)  */
    uint64_t pos;             /* absolute output-word position             */
    uint32_t b0, b1, b2, b3;  /* block pos/4                               */
} oa_philox_stream;

OA_FN void oa_philox_fill(oa_philox_stream* s, uint64_t block)
{
    oa_u32x4 c;
    c.v[0] = (uint32_t)block;
    c.v[1] = (uint32_t)(block >> 32);
    c.v[2] = (uint32_t)s->photon;
    c.v[3] = (uint32_t)(s->photon >> 32);
    oa_u32x4 o = oa_philox4x32_10(c, (uint32_t)s->key, (uint32_t)(s->key >> 32));
    s->b0 = o.v[0]; s->b1 = o.v[1]; s->b2 = o.v[2]; s->b3 = o.v[3];
}

/*
 * oa_rng_init: stream for absolute photon index `photon`, advanced by
 * `offset + event_idx*skipahead_event_offset` words, mirroring the
 * subsequence/skipahead structure of qudarap/qrng.h:131-137.
 */
OA_FN void oa_rng_init(oa_philox_stream* s, uint64_t seed, uint64_t offset,
                       uint64_t skipahead_event_offset, uint64_t event_idx,
                       uint64_t photon)
{
    s->key = seed;
    s->photon = photon;
    s->pos = offset + skipahead_event_offset * event_idx;
    oa_philox_fill(s, s->pos >> 2);
    const uint32_t skip = (uint32_t)(s->pos & 3u);     /* start inside the block */
    for (uint32_t i = 0; i < 3u; ++i) {
        if (i < skip) { s->b0 = s->b1; s->b1 = s->b2; s->b2 = s->b3; }
    }
}

OA_FN uint32_t oa_rng_next_u32(oa_philox_stream* s)
{
    uint32_t x = s->b0;
    s->b0 = s->b1; s->b1 = s->b2; s->b2 = s->b3;
    s->pos += 1u;
    if ((s->pos & 3u) == 0u) oa_philox_fill(s, s->pos >> 2);
    return x;
}

OA_FN float oa_rng_uniform(oa_philox_stream* s)
{
    return oa_u32_to_uniform(oa_rng_next_u32(s));
}

/* --------------------------------------------------------------- XORWOW */

/*
 * xorwow: G. Marsaglia, "Xorshift RNGs", J. Stat. Softw. 8(14), 2003, section
 * 3.1: a 160-bit xorshift (x0..x4, shifts 2, 1, 4) plus a Weyl sequence d with
 * increment 362437; output word = d + x4 after the update. The reference's
 * qrng<XORWOW> (qudarap/qrng.h:68-106) draws from a generator with this
 * recurrence. Seeding and subsequences are this project's own convention,
 * not curand's (whose seeding scramble and skipahead tables are not
 * reproduced):
 *   subsequence (photon) start: three outputs w0, w1, w2 of SplitMix64
 *     (G. Steele, D. Lea, C. Flood, "Fast splittable pseudorandom number
 *     generators", OOPSLA 2014) whose state starts at mix(seed + gamma) XOR
 *     photon; x0 = lo(w0), x1 = hi(w0), x2 = lo(w1), x3 = hi(w1),
 *     x4 = lo(w2), d = hi(w2) (x0 = 1 if x0..x4 would all be zero).
 *     Distinct photons (< 2^63) start from distinct SplitMix64 inputs.
 *   skipahead by n words (offset + event * skipahead_event_offset, as
 *     qrng.h:84-89 skips): the xorshift part is linear over GF(2), so its
 *     state is multiplied by J = M^n (M the 160x160 one-step matrix), and
 *     d += 362437 n (mod 2^32). J is formed on the host by square and
 *     multiply; oa_xorwow_apply multiplies a state by it.
 * Uniform mapping as for Philox: u = ((x >> 8) + 1) 2^-24 in (0, 1].
 */
typedef struct {
    uint32_t x0, x1, x2, x3, x4, d;
} oa_xorwow_state;

#define OA_XORWOW_WEYL 362437u
#define OA_SPLITMIX_GAMMA 0x9E3779B97F4A7C15ull
#define OA_XORWOW_JUMP_WORDS 800      /* 160 columns x 5 words */

OA_FN uint32_t oa_xorwow_next(oa_xorwow_state* s)
{
    const uint32_t t = s->x0 ^ (s->x0 >> 2);
    s->x0 = s->x1; s->x1 = s->x2; s->x2 = s->x3; s->x3 = s->x4;
    s->x4 = (s->x4 ^ (s->x4 << 4)) ^ (t ^ (t << 1));
    s->d += OA_XORWOW_WEYL;
    return s->d + s->x4;
}

OA_FN float oa_xorwow_uniform(oa_xorwow_state* s)
{
    return oa_u32_to_uniform(oa_xorwow_next(s));
}

OA_FN uint64_t oa_splitmix64_mix(uint64_t z)
{
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

OA_FN void oa_xorwow_seed(oa_xorwow_state* s, uint64_t seed, uint64_t photon)
{
    uint64_t sm = oa_splitmix64_mix(seed + OA_SPLITMIX_GAMMA) ^ photon;
    sm += OA_SPLITMIX_GAMMA; const uint64_t w0 = oa_splitmix64_mix(sm);
    sm += OA_SPLITMIX_GAMMA; const uint64_t w1 = oa_splitmix64_mix(sm);
    sm += OA_SPLITMIX_GAMMA; const uint64_t w2 = oa_splitmix64_mix(sm);
    s->x0 = (uint32_t)w0; s->x1 = (uint32_t)(w0 >> 32);
    s->x2 = (uint32_t)w1; s->x3 = (uint32_t)(w1 >> 32);
    s->x4 = (uint32_t)w2; s->d = (uint32_t)(w2 >> 32);
    if ((s->x0 | s->x1 | s->x2 | s->x3 | s->x4) == 0u) s->x0 = 1u;
}

/* state x <- J x for a jump matrix J stored as 160 columns of 5 words:
   column j (words 5j..5j+4) is the image of bit (j % 32) of word j / 32 */
#define OA_XORWOW_COLS(v, first)                                                          \
    for (int b = 0; b < 32; ++b) {                                                        \
        const uint32_t m = 0u - (((v) >> b) & 1u);                                        \
        const uint32_t* c = J + 5 * ((first) + b);                                        \
        o0 ^= c[0] & m; o1 ^= c[1] & m; o2 ^= c[2] & m; o3 ^= c[3] & m; o4 ^= c[4] & m;   \
    }

OA_FN void oa_xorwow_apply(oa_xorwow_state* s, const uint32_t* J)
{
    uint32_t o0 = 0u, o1 = 0u, o2 = 0u, o3 = 0u, o4 = 0u;
    OA_XORWOW_COLS(s->x0, 0)
    OA_XORWOW_COLS(s->x1, 32)
    OA_XORWOW_COLS(s->x2, 64)
    OA_XORWOW_COLS(s->x3, 96)
    OA_XORWOW_COLS(s->x4, 128)
    s->x0 = o0; s->x1 = o1; s->x2 = o2; s->x3 = o3; s->x4 = o4;
}

/* skip n words: J = M^n (precomputed for this n, or NULL when n == 0) */
OA_FN void oa_xorwow_skip(oa_xorwow_state* s, const uint32_t* J, uint64_t n)
{
    if (n != 0u && J) oa_xorwow_apply(s, J);
    s->d += OA_XORWOW_WEYL * (uint32_t)n;
}

/*
 * Host-side jump matrices (plain C, shared by the runtime and the checker).
 * A matrix is 160 columns x 5 words. Products use a table of the XORs of
 * every 8-column group of the left factor (the "four Russians" method), so
 * one product costs 20 x 256 + 160 x 20 five-word XORs.
 */
OA_FN void oa_xorwow_step_matrix(uint32_t* M)
{
    for (int j = 0; j < 160; ++j) {
        oa_xorwow_state s;
        uint32_t in[5] = { 0u, 0u, 0u, 0u, 0u };
        in[j / 32] = 1u << (j % 32);
        s.x0 = in[0]; s.x1 = in[1]; s.x2 = in[2]; s.x3 = in[3]; s.x4 = in[4]; s.d = 0u;
        (void)oa_xorwow_next(&s);
        uint32_t* c = M + 5 * j;
        c[0] = s.x0; c[1] = s.x1; c[2] = s.x2; c[3] = s.x3; c[4] = s.x4;
    }
}

/* C = A B (C may alias neither A nor B); tab: 20*256*5 words of workspace */
OA_FN void oa_xorwow_matmul(uint32_t* C, const uint32_t* A, const uint32_t* B, uint32_t* tab)
{
    for (int g = 0; g < 20; ++g) {
        uint32_t* t = tab + 5 * 256 * g;
        for (int k = 0; k < 5; ++k) t[k] = 0u;
        for (int v = 1; v < 256; ++v) {
            const int low = v & -v;                      /* lowest set bit of v */
            int bit = 0;
            while ((1 << bit) != low) ++bit;
            const uint32_t* a = A + 5 * (8 * g + bit);
            const uint32_t* prev = t + 5 * (v ^ low);
            for (int k = 0; k < 5; ++k) t[5 * v + k] = prev[k] ^ a[k];
        }
    }
    for (int j = 0; j < 160; ++j) {
        const uint32_t* b = B + 5 * j;
        uint32_t o[5] = { 0u, 0u, 0u, 0u, 0u };
        for (int g = 0; g < 20; ++g) {
            const uint32_t v = (b[g / 4] >> (8 * (g % 4))) & 0xffu;
            const uint32_t* t = tab + 5 * (256 * g + v);
            for (int k = 0; k < 5; ++k) o[k] ^= t[k];
        }
        for (int k = 0; k < 5; ++k) C[5 * j + k] = o[k];
    }
}

/* J = M^n by square and multiply; work: 3 * 800 + 20*256*5 words */
OA_FN void oa_xorwow_jump_matrix(uint64_t n, uint32_t* J, uint32_t* work)
{
    uint32_t* P = work;            /* M^(2^k)  */
    uint32_t* T = work + 800;      /* product  */
    uint32_t* tab = work + 1600;
    oa_xorwow_step_matrix(P);
    for (int j = 0; j < 160; ++j)
        for (int k = 0; k < 5; ++k) J[5 * j + k] = (k == j / 32) ? (1u << (j % 32)) : 0u;
    while (n) {
        if (n & 1u) {
            oa_xorwow_matmul(T, P, J, tab);
            for (int i = 0; i < 800; ++i) J[i] = T[i];
        }
        n >>= 1;
        if (n) {
            oa_xorwow_matmul(T, P, P, tab);
            for (int i = 0; i < 800; ++i) P[i] = T[i];
        }
    }
}
#define OA_XORWOW_WORK_WORDS (1600 + 20 * 256 * 5)

#endif /* OA_PORTABLE_H */
