// dq_device.h -- device-side building blocks shared by the engine's kernels:
// complex helpers, the Philox4x64-10 noise stream with its Box-Muller
// normals, and wave64 reductions.  gfx950 only.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace dq {

constexpr int kWave = 64;

struct cplx { double re, im; };

__device__ __forceinline__ double2 ld2(const double2* p) { return *p; }

// ---------------------------------------------------------------------------
// Philox4x64-10 (Salmon et al., SC'11).  The reference's MRG32k3a states are
// shared by `elem % 4096` (gpuSysHelper.cu:50,94) and time-seeded; a counter
// keyed by (seed, trajectory) and counted by (site, step, stream, substep/4)
// needs no state at all.
// ---------------------------------------------------------------------------
__host__ __device__ __forceinline__ void philox4x64_10(uint64_t c[4], uint64_t k0, uint64_t k1)
{
    const uint64_t M0 = 0xD2E7470EE14C6C93ULL, M1 = 0xCA5A826395121157ULL;
    const uint64_t W0 = 0x9E3779B97F4A7C15ULL, W1 = 0xBB67AE8584CAA73BULL;
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        if (r > 0) { k0 += W0; k1 += W1; }
#if defined(__HIP_DEVICE_COMPILE__)
        uint64_t hi0 = __umul64hi(M0, c[0]), lo0 = M0 * c[0];
        uint64_t hi1 = __umul64hi(M1, c[2]), lo1 = M1 * c[2];
#else
        unsigned __int128 p0 = (unsigned __int128)M0 * c[0];
        unsigned __int128 p1 = (unsigned __int128)M1 * c[2];
        uint64_t hi0 = (uint64_t)(p0 >> 64), lo0 = (uint64_t)p0;
        uint64_t hi1 = (uint64_t)(p1 >> 64), lo1 = (uint64_t)p1;
#endif
        uint64_t n0 = hi1 ^ c[1] ^ k0;
        uint64_t n2 = hi0 ^ c[3] ^ k1;
        c[0] = n0; c[1] = lo1; c[2] = n2; c[3] = lo0;
    }
}

__host__ __device__ __forceinline__ double bits_to_f64(uint64_t b) { return __builtin_bit_cast(double, b); }
__host__ __device__ __forceinline__ uint64_t f64_to_bits(double d) { return __builtin_bit_cast(uint64_t, d); }

// The three functions below are built from +, *, fma and bit operations
// only, evaluated with contraction off, so that every platform that rounds
// those operations per IEEE 754 produces the same bits.
__host__ __device__ inline double rng_log(double x)
{
#pragma clang fp contract(off)
    // x = 2^e f, f in [sqrt(1/2), sqrt(2));  ln f = 2 atanh(s), s = (f-1)/(f+1)
    uint64_t b = f64_to_bits(x);
    int e = (int)((b >> 52) & 0x7ff) - 1023;
    double f = bits_to_f64((b & 0x000fffffffffffffULL) | 0x3ff0000000000000ULL);
    if (f > 0x1.6a09e667f3bcdp+0) { f = f * 0.5; e = e + 1; }
    double num = f - 1.0;
    double den = f + 1.0;
    double r = 1.0 - 0.2426 * den;              // 1/den to 4%, den in [1.70, 2.42]
    for (int it = 0; it < 6; ++it) {
        double er = __builtin_fma(-den, r, 1.0);
        r = __builtin_fma(r, er, r);
    }
    double s = num * r;
    double res = __builtin_fma(-den, s, num);
    s = __builtin_fma(res, r, s);
    double z = s * s;
    double p = 0x1.47ae147ae147bp-5;            // 1/25 ... 1/3 (Horner in s^2)
    p = __builtin_fma(p, z, 0x1.642c8590b2164p-5);
    p = __builtin_fma(p, z, 0x1.8618618618618p-5);
    p = __builtin_fma(p, z, 0x1.af286bca1af28p-5);
    p = __builtin_fma(p, z, 0x1.e1e1e1e1e1e1ep-5);
    p = __builtin_fma(p, z, 0x1.1111111111111p-4);
    p = __builtin_fma(p, z, 0x1.3b13b13b13b14p-4);
    p = __builtin_fma(p, z, 0x1.745d1745d1746p-4);
    p = __builtin_fma(p, z, 0x1.c71c71c71c71cp-4);
    p = __builtin_fma(p, z, 0x1.2492492492492p-3);
    p = __builtin_fma(p, z, 0x1.999999999999ap-3);
    p = __builtin_fma(p, z, 0x1.5555555555555p-2);
    double t = (s * z) * p;
    double lnf = __builtin_fma(2.0, s, 2.0 * t);
    double de = (double)e;
    return __builtin_fma(de, 0x1.62e42fefa39efp-1, __builtin_fma(de, 0x1.abc9e3b39803fp-56, lnf));
}

__host__ __device__ inline double rng_sqrt(double x)
{
#pragma clang fp contract(off)
    if (!(x > 0.0)) return 0.0;
    uint64_t b = f64_to_bits(x);
    double y = bits_to_f64(0x5fe6eb50c7b537a9ULL - (b >> 1));   // rsqrt seed
    double hx = 0.5 * x;
    for (int it = 0; it < 5; ++it) {
        double yy = y * y;
        double t = __builtin_fma(-hx, yy, 1.5);
        y = y * t;
    }
    double s = x * y;
    double r = __builtin_fma(-s, s, x);
    return __builtin_fma(r, 0.5 * y, s);
}

__host__ __device__ inline void rng_sincos_2pi(double u, double* s_out, double* c_out)
{
#pragma clang fp contract(off)
    double v = 4.0 * u;
    double q = __builtin_floor(v + 0.5);
    double t = v - q;
    double ph = t * 0x1.921fb54442d18p+0;
    double z = ph * ph;
    double ps = -0x1.2f49b46814157p-57;        // Taylor coefficients (-1)^k/(2k+1)!
    ps = __builtin_fma(ps, z, 0x1.952c77030ad4ap-49);
    ps = __builtin_fma(ps, z, -0x1.ae7f3e733b81fp-41);
    ps = __builtin_fma(ps, z, 0x1.6124613a86d09p-33);
    ps = __builtin_fma(ps, z, -0x1.ae64567f544e4p-26);
    ps = __builtin_fma(ps, z, 0x1.71de3a556c734p-19);
    ps = __builtin_fma(ps, z, -0x1.a01a01a01a01ap-13);
    ps = __builtin_fma(ps, z, 0x1.1111111111111p-7);
    ps = __builtin_fma(ps, z, -0x1.5555555555555p-3);
    double sn = __builtin_fma(ph * z, ps, ph);
    double pc = 0x1.e542ba4020225p-62;         // (-1)^k/(2k)!
    pc = __builtin_fma(pc, z, -0x1.6827863b97d97p-53);
    pc = __builtin_fma(pc, z, 0x1.ae7f3e733b81fp-45);
    pc = __builtin_fma(pc, z, -0x1.93974a8c07c9dp-37);
    pc = __builtin_fma(pc, z, 0x1.1eed8eff8d898p-29);
    pc = __builtin_fma(pc, z, -0x1.27e4fb7789f5cp-22);
    pc = __builtin_fma(pc, z, 0x1.a01a01a01a01ap-16);
    pc = __builtin_fma(pc, z, -0x1.6c16c16c16c17p-10);
    pc = __builtin_fma(pc, z, 0x1.5555555555555p-5);
    pc = __builtin_fma(pc, z, -0.5);
    double cs = __builtin_fma(z, pc, 1.0);
    int qi = ((int)q) & 3;
    double so, co;
    if (qi == 0) { so = sn; co = cs; }
    else if (qi == 1) { so = cs; co = -sn; }
    else if (qi == 2) { so = -sn; co = -cs; }
    else { so = -cs; co = sn; }
    *s_out = so;
    *c_out = co;
}

__host__ __device__ inline void box_muller(uint64_t a, uint64_t b, double* z0, double* z1)
{
#pragma clang fp contract(off)
    double u1 = (double)((a >> 11) + 1) * 0x1p-53;
    double u2 = (double)(b >> 11) * 0x1p-53;
    double r = rng_sqrt(-2.0 * rng_log(u1));
    double s, c;
    rng_sincos_2pi(u2, &s, &c);
    *z0 = r * c;
    *z1 = r * s;
}

enum : uint64_t { kStreamStatic = 1, kStreamOuInit = 2, kStreamOu = 3, kStreamHsr = 4, kStreamZz = 5 };

// uniform phases u in [0,1) of the Zhong-Zhao oscillators 4q..4q+3 of a site
// (fixed per trajectory: the counter has no step)
__host__ __device__ inline void zz_phase4(uint64_t seed, uint64_t traj, uint64_t site, uint64_t q, double u[4])
{
    uint64_t c[4] = {site, q, kStreamZz, 0};
    philox4x64_10(c, seed, traj);
    for (int r = 0; r < 4; ++r) u[r] = (double)(c[r] >> 11) * 0x1p-53;
}

// cos(2 pi (u + w t)) with the argument reduced to [0,1)
__host__ __device__ inline double zz_cos(double u, double w2pi, double t)
{
#pragma clang fp contract(off)
    double x = u + w2pi * t;
    x = x - __builtin_floor(x);
    if (x >= 1.0) x = 0.0;
    double s, c;
    rng_sincos_2pi(x, &s, &c);
    return c;
}

// four normals of counter (site, step, stream, block)
__host__ __device__ inline void normal4(uint64_t seed, uint64_t traj, uint64_t site, uint64_t step,
                                        uint64_t stream, uint64_t block, double z[4])
{
    uint64_t c[4] = {site, step, stream, block};
    philox4x64_10(c, seed, traj);
    box_muller(c[0], c[1], &z[0], &z[1]);
    box_muller(c[2], c[3], &z[2], &z[3]);
}

// ---------------------------------------------------------------------------
// wave64 reductions
// ---------------------------------------------------------------------------
__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
    return v;
}

// sum over groups of `width` consecutive lanes (width power of two <= 64);
// every lane of a group ends with the group's sum
__device__ __forceinline__ double group_sum(double v, int width)
{
    for (int off = width >> 1; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
    return v;
}

__device__ __forceinline__ int uniform(int v) { return __builtin_amdgcn_readfirstlane(v); }

// blockIdx -> logical tile so that consecutive logical tiles land on the same
// XCD (MI355X dispatches workgroups round-robin over its 8 XCDs; neighbouring
// rows then share that XCD's L2).  `padded` is a multiple of 8.
__device__ __forceinline__ int xcd_tile(int bid, int padded)
{
    int per = padded >> 3;
    return (bid & 7) * per + (bid >> 3);
}

}  // namespace dq
