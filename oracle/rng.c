/*
 * rng.c -- oracle restatement of the noise source.  TEST INFRASTRUCTURE.
 *
 * The reference draws its Kubo-Anderson increments from per-thread
 * MRG32k3a states seeded from the wall clock (gpuSysHelper.cu:73-128,
 * systemClass.cu:874-905) and its static disorder / initial offsets from
 * Boost mt19937 (systemClass.cu:759-804, 1028-1036).  Neither stream can be
 * reproduced, so the method is restated on a counter-based generator:
 * Philox4x64-10 (Salmon, Moraes, Dror, Shaw, SC'11), keyed by
 * (seed, trajectory) and counted by (site, step, stream, substep/4).
 * Normals come from the Box-Muller transform with a logarithm, a square
 * root and sin/cos(2 pi u) built here from +, *, fma and integer bit
 * operations only, so that the HIP engine, which restates the same
 * operation sequence, produces bit-identical normals.
 *
 * Compile with -ffp-contract=off: every rounding below is explicit.
 */
#include "dq_oracle.h"
#include <math.h>
#include <string.h>

#define PH_M0 0xD2E7470EE14C6C93ULL
#define PH_M1 0xCA5A826395121157ULL
#define PH_W0 0x9E3779B97F4A7C15ULL
#define PH_W1 0xBB67AE8584CAA73BULL

void orc_philox4x64_10(const uint64_t ctr[4], const uint64_t key[2], uint64_t out[4])
{
    uint64_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint64_t k0 = key[0], k1 = key[1];
    for (int r = 0; r < 10; ++r) {
        if (r > 0) { k0 += PH_W0; k1 += PH_W1; }
        unsigned __int128 p0 = (unsigned __int128)PH_M0 * c0;
        unsigned __int128 p1 = (unsigned __int128)PH_M1 * c2;
        uint64_t hi0 = (uint64_t)(p0 >> 64), lo0 = (uint64_t)p0;
        uint64_t hi1 = (uint64_t)(p1 >> 64), lo1 = (uint64_t)p1;
        uint64_t n0 = hi1 ^ c1 ^ k0;
        uint64_t n2 = hi0 ^ c3 ^ k1;
        c0 = n0; c1 = lo1; c2 = n2; c3 = lo0;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

static double bits2d(uint64_t b) { double d; memcpy(&d, &b, 8); return d; }
static uint64_t d2bits(double d) { uint64_t b; memcpy(&b, &d, 8); return b; }

/* 1/den for den in [1.70, 2.42] (f + 1 below): linear seed within 4%,
 * six Newton steps (quadratic convergence reaches rounding after four). */
static double orc_recip_14(double den)
{
    double r = 1.0 - 0.2426 * den;
    for (int it = 0; it < 6; ++it) {
        double e = fma(-den, r, 1.0);
        r = fma(r, e, r);
    }
    return r;
}

/* natural log for normal positive x:  x = 2^e f, f in [sqrt(1/2), sqrt(2)),
 * ln f = 2 atanh(s), s = (f-1)/(f+1), series in s^2 to s^23. */
double orc_log(double x)
{
    uint64_t b = d2bits(x);
    int e = (int)((b >> 52) & 0x7ff) - 1023;
    double f = bits2d((b & 0x000fffffffffffffULL) | 0x3ff0000000000000ULL);
    if (f > 0x1.6a09e667f3bcdp+0) { f = f * 0.5; e = e + 1; }
    double num = f - 1.0;                   /* exact (Sterbenz) */
    double den = f + 1.0;
    double r = orc_recip_14(den);
    double s = num * r;
    double res = fma(-den, s, num);         /* correct the quotient */
    s = fma(res, r, s);
    double z = s * s;
    double p = 0x1.47ae147ae147bp-5;        /* 1/25 */
    p = fma(p, z, 0x1.642c8590b2164p-5);    /* 1/23 */
    p = fma(p, z, 0x1.8618618618618p-5);    /* 1/21 */
    p = fma(p, z, 0x1.af286bca1af28p-5);    /* 1/19 */
    p = fma(p, z, 0x1.e1e1e1e1e1e1ep-5);    /* 1/17 */
    p = fma(p, z, 0x1.1111111111111p-4);    /* 1/15 */
    p = fma(p, z, 0x1.3b13b13b13b14p-4);    /* 1/13 */
    p = fma(p, z, 0x1.745d1745d1746p-4);    /* 1/11 */
    p = fma(p, z, 0x1.c71c71c71c71cp-4);    /* 1/9 */
    p = fma(p, z, 0x1.2492492492492p-3);    /* 1/7 */
    p = fma(p, z, 0x1.999999999999ap-3);    /* 1/5 */
    p = fma(p, z, 0x1.5555555555555p-2);    /* 1/3 */
    double t = (s * z) * p;                 /* s^3 (1/3 + ...) */
    double lnf = fma(2.0, s, 2.0 * t);
    double de = (double)e;
    return fma(de, 0x1.62e42fefa39efp-1, fma(de, 0x1.abc9e3b39803fp-56, lnf));
}

/* sqrt via reciprocal-sqrt Newton iteration and one residual correction */
double orc_sqrt(double x)
{
    if (!(x > 0.0)) return 0.0;
    uint64_t b = d2bits(x);
    double y = bits2d(0x5fe6eb50c7b537a9ULL - (b >> 1));
    double hx = 0.5 * x;
    for (int it = 0; it < 5; ++it) {
        double yy = y * y;
        double t = fma(-hx, yy, 1.5);
        y = y * t;
    }
    double s = x * y;
    double r = fma(-s, s, x);
    return fma(r, 0.5 * y, s);
}

/* sin and cos of 2 pi u for u in [0, 1): quadrant from 4u (exact), then
 * Taylor polynomials of sin/cos on |phi| <= pi/4. */
void orc_sincos_2pi(double u, double *s_out, double *c_out)
{
    double v = 4.0 * u;                     /* exact */
    double q = floor(v + 0.5);              /* exact: v has <= 53 bits */
    double t = v - q;                       /* exact, |t| <= 1/2 */
    double ph = t * 0x1.921fb54442d18p+0;   /* (pi/2) t */
    double z = ph * ph;
    double ps = -0x1.2f49b46814157p-57;     /* -1/19! */
    ps = fma(ps, z, 0x1.952c77030ad4ap-49);
    ps = fma(ps, z, -0x1.ae7f3e733b81fp-41);
    ps = fma(ps, z, 0x1.6124613a86d09p-33);
    ps = fma(ps, z, -0x1.ae64567f544e4p-26);
    ps = fma(ps, z, 0x1.71de3a556c734p-19);
    ps = fma(ps, z, -0x1.a01a01a01a01ap-13);
    ps = fma(ps, z, 0x1.1111111111111p-7);
    ps = fma(ps, z, -0x1.5555555555555p-3);
    double sn = fma(ph * z, ps, ph);
    double pc = 0x1.e542ba4020225p-62;      /* 1/20! */
    pc = fma(pc, z, -0x1.6827863b97d97p-53);
    pc = fma(pc, z, 0x1.ae7f3e733b81fp-45);
    pc = fma(pc, z, -0x1.93974a8c07c9dp-37);
    pc = fma(pc, z, 0x1.1eed8eff8d898p-29);
    pc = fma(pc, z, -0x1.27e4fb7789f5cp-22);
    pc = fma(pc, z, 0x1.a01a01a01a01ap-16);
    pc = fma(pc, z, -0x1.6c16c16c16c17p-10);
    pc = fma(pc, z, 0x1.5555555555555p-5);
    pc = fma(pc, z, -0.5);
    double cs = fma(z, pc, 1.0);
    int qi = ((int)q) & 3;
    switch (qi) {
    case 0: *s_out = sn;  *c_out = cs;  break;
    case 1: *s_out = cs;  *c_out = -sn; break;
    case 2: *s_out = -sn; *c_out = -cs; break;
    default: *s_out = -cs; *c_out = sn; break;
    }
}

static void box_muller(uint64_t a, uint64_t b, double *z0, double *z1)
{
    double u1 = (double)((a >> 11) + 1) * 0x1p-53;   /* (0, 1] */
    double u2 = (double)(b >> 11) * 0x1p-53;         /* [0, 1) */
    double r = orc_sqrt(-2.0 * orc_log(u1));
    double s, c;
    orc_sincos_2pi(u2, &s, &c);
    *z0 = r * c;
    *z1 = r * s;
}

void orc_normal4(uint64_t seed, uint64_t traj, uint64_t c0, uint64_t c1,
                 uint64_t c2, uint64_t c3, double z[4])
{
    uint64_t ctr[4] = {c0, c1, c2, c3}, key[2] = {seed, traj}, o[4];
    orc_philox4x64_10(ctr, key, o);
    box_muller(o[0], o[1], &z[0], &z[1]);
    box_muller(o[2], o[3], &z[2], &z[3]);
}

double orc_normal(uint64_t seed, uint64_t traj, uint64_t site, uint64_t step,
                  uint64_t stream, uint64_t sub)
{
    double z[4];
    orc_normal4(seed, traj, site, step, stream, sub >> 2, z);
    return z[sub & 3];
}
