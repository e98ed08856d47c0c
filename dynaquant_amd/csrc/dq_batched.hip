// dq_batched.hip -- batched (B >= 64 padded) kernels of the Krylov cycle.
//
// SpMM: one wave per block of kRB rows x 64-128 trajectories over the union
// of the rows' columns (register reuse of every gathered x run), with the
// sparse structure read through wave-uniform scalar loads.
//
// Update and combine: a wave covers 16 trajectories x 4 groups of Krylov
// vectors (lane = 16 * kgroup + t).  Each lane streams only the vectors of
// its group, the four partial sums of a row are combined with two xor
// shuffles (no barrier), and each lane keeps FP64 Gram accumulators only
// for its own group's vectors -- a quarter of the registers of a
// one-lane-per-trajectory layout, so several waves per SIMD stay resident.
#include "dq_device.h"
#include "dq_internal.h"
#include "dq_reduce.h"

namespace dq {

template <typename VT> struct BVec2;
template <> struct BVec2<float> { using T = float2; };
template <> struct BVec2<double> { using T = double2; };

static inline int bpad8(int x) { return (x + 7) & ~7; }

__device__ __forceinline__ float bcast(float v, int l)
{
    return __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), l));
}
__device__ __forceinline__ double bcast(double v, int l)
{
    const long long x = __builtin_bit_cast(long long, v);
    const int lo = __builtin_amdgcn_readlane((int)(x & 0xffffffff), l);
    const int hi = __builtin_amdgcn_readlane((int)(x >> 32), l);
    return __builtin_bit_cast(double, ((long long)hi << 32) | (unsigned int)lo);
}

// block partials for batched lanes (lane = trajectory of a 64-wide tile)
template <int NQ>
__device__ __forceinline__ void bstore_partials(const double (&acc)[NQ], int nq, double* part, int ntile,
                                                int tile, int Bp, int b)
{
    __shared__ double lds[4][4][64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int q0 = 0; q0 < NQ; q0 += 4) {
        if (q0 < nq) {
            __syncthreads();
#pragma unroll
            for (int u = 0; u < 4; ++u)
                if (q0 + u < NQ) lds[u][wave][lane] = acc[q0 + u < NQ ? q0 + u : 0];
            __syncthreads();
            if (wave == 0) {
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const int q = q0 + u;
                    if (q < nq && q < NQ)
                        part[((size_t)q * ntile + tile) * Bp + b] =
                            (lds[u][0][lane] + lds[u][1][lane]) + (lds[u][2][lane] + lds[u][3][lane]);
                }
            }
        }
    }
}

// ---------------------------------------------------------------------------
// SpMM  w = d .* x + g * J x  (lanczosClass.cu:618-661 per trajectory), with
// the FP64 couplings and FP64 sums also for FP32 vectors (x widened exactly,
// w rounded once): FP32 couplings are a fixed perturbation of H, and FP32 sums
// over a row's ~60 couplings, whose partial sums cancel to a much smaller w,
// were the largest part of the mixed-precision population error.
// Row-blocked: a wave owns kRB consecutive rows (2, 4 or 8, chosen by the
// autotuner) and walks the union of their columns, so each gathered x run (a
// contiguous 512 B / 1 KB segment of trajectories) feeds kRB rows from
// registers.  For the N = 12,000 dipole tube kRB = 4 cuts the gathers per row
// from 62 to 21 at 1.33x the FMAs of plain CSR; kRB = 8 gives 14 at 1.78x.
// ---------------------------------------------------------------------------
template <typename VT> struct Vec4;
template <> struct Vec4<float> { using T = float4; };

// explicit fused multiply-adds, so that every instantiation (trajectories per
// lane, rows per wave) rounds each trajectory's sums identically
__device__ __forceinline__ float fmav(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ __forceinline__ double fmav(double a, double b, double c) { return __builtin_fma(a, b, c); }

// Each wave owns one row block (kRB rows) x 64*TPL trajectories; the four
// waves of a workgroup take consecutive row blocks.  Column offsets and the
// kRB values of each union column are wave-uniform scalar loads, so a gather
// costs one vector memory instruction addressed as uniform base + lane offset
// and the vector ALU only runs the FMAs.  TPL trajectories per lane: 2 in FP32
// (one 16-byte load per lane), 1 in FP64.
// G gathers of x rows rc[p..p+G-1] (TPL trajectories per lane)
template <typename VT, int TPL, int G, typename V2>
__device__ __forceinline__ void gatherg(const V2* __restrict__ xs, const int32_t* __restrict__ rc, int p, int Bp,
                                        int lane, VT (&xr)[G][TPL], VT (&xi)[G][TPL])
{
#pragma unroll
    for (int u = 0; u < G; ++u) {
        const V2* xc = xs + (size_t)uniform(rc[p + u]) * Bp;
        if constexpr (TPL == 2) {
            const float4 q = *reinterpret_cast<const float4*>(xc + 2 * lane);
            xr[u][0] = q.x; xi[u][0] = q.y; xr[u][1] = q.z; xi[u][1] = q.w;
        } else {
            const V2 q = xc[lane];
            xr[u][0] = q.x; xi[u][0] = q.y;
        }
    }
}

// rows r of the block += value(p+u, r) * x(p+u), in column order, in FP64
// (FP64 values; an FP32 x is widened exactly)
template <typename VT, int TPL, int kRB, int G>
__device__ __forceinline__ void fmag(const double* __restrict__ rv, int p, const VT (&xr)[G][TPL],
                                     const VT (&xi)[G][TPL], double (&ar)[TPL][kRB], double (&ai)[TPL][kRB])
{
#pragma unroll
    for (int u = 0; u < G; ++u) {
#pragma unroll
        for (int r = 0; r < kRB; ++r) {
            const double v = rv[(size_t)(p + u) * kRB + r];
#pragma unroll
            for (int t = 0; t < TPL; ++t) {
                ar[t][r] = __builtin_fma(v, (double)xr[u][t], ar[t][r]);
                ai[t][r] = __builtin_fma(v, (double)xi[u][t], ai[t][r]);
            }
        }
    }
}

template <typename VT, int TPL, int kRB, bool PIPE>
__global__ __launch_bounds__(kThreads) void k_spmm_b(Dev d, int j, int last, int tt0, int ntt)
{
    using V2 = typename BVec2<VT>::T;
    constexpr int TW = 64 * TPL;                         // trajectories per wave
    const int logical = xcd_tile(blockIdx.x, gridDim.x);
    if (logical >= d.ntile_s * ntt) return;
    const int rtile = logical / ntt, ttile = tt0 + logical % ntt;
    const int lane = threadIdx.x & 63;
    const int wave = uniform(threadIdx.x >> 6);
    const int Bp = d.Bp;
    const int b0 = ttile * TW + lane * TPL;              // first trajectory of this lane
    const size_t nB = (size_t)d.n * Bp;
    const V2* __restrict__ xs = reinterpret_cast<const V2*>(d.V) + (size_t)j * d.vs + (size_t)ttile * TW;
    V2* __restrict__ y = reinterpret_cast<V2*>(d.V) + (size_t)(j + 1) * d.vs;
    const V2* __restrict__ x = reinterpret_cast<const V2*>(d.V) + (size_t)j * d.vs;
    const VT* __restrict__ diag = reinterpret_cast<const VT*>(d.diag);
    const double* __restrict__ rv = d.rb_val64;
    const int32_t* __restrict__ rc = d.rb_col;
    const int nblk = (d.n + kRB - 1) / kRB;
    const int rb = rtile * 4 + wave;
    double acc[TPL][2];
#pragma unroll
    for (int t = 0; t < TPL; ++t) { acc[t][0] = 0.0; acc[t][1] = 0.0; }
    // slots are sorted by n_active: a row block at or beyond the slice's first
    // n_active is inactive for every lane (its w rows stay 0)
    const int slice_nact = uniform(d.nact[ttile * TW]);
    if (rb < nblk && rb * kRB < slice_nact) {
        const int u0 = uniform(d.rb_ptr[rb]), u1 = uniform(d.rb_ptr[rb + 1]);
        double ar[TPL][kRB], ai[TPL][kRB];
#pragma unroll
        for (int t = 0; t < TPL; ++t)
#pragma unroll
            for (int r = 0; r < kRB; ++r) { ar[t][r] = 0.0; ai[t][r] = 0.0; }
        int p = u0;
        if constexpr (PIPE) {
            // two 4-column register buffers (the registers of one 8-column
            // buffer): the gathers of the next 4 columns are in flight while
            // the current 4 are multiplied in
            const int nb = (u1 - p) / 4;
            VT xar[4][TPL], xai[4][TPL], xbr[4][TPL], xbi[4][TPL];
            if (nb > 0) gatherg<VT, TPL, 4>(xs, rc, p, Bp, lane, xar, xai);
            int bb = 0;
            for (; bb + 2 <= nb; bb += 2) {
                gatherg<VT, TPL, 4>(xs, rc, p + 4, Bp, lane, xbr, xbi);
                fmag<VT, TPL, kRB, 4>(rv, p, xar, xai, ar, ai);
                p += 4;
                if (bb + 2 < nb) gatherg<VT, TPL, 4>(xs, rc, p + 4, Bp, lane, xar, xai);
                fmag<VT, TPL, kRB, 4>(rv, p, xbr, xbi, ar, ai);
                p += 4;
            }
            if (bb < nb) {
                fmag<VT, TPL, kRB, 4>(rv, p, xar, xai, ar, ai);
                p += 4;
            }
        } else {
            // the column offsets of the next 8 columns are loaded while this
            // group's gathers are in flight, so a scalar-load round trip no
            // longer precedes every group of gathers
            int cur[8], nxt[8];
            if (p + 8 <= u1) {
#pragma unroll
                for (int u = 0; u < 8; ++u) cur[u] = uniform(rc[p + u]);
            }
            for (; p + 8 <= u1; p += 8) {
                VT xr[8][TPL], xi[8][TPL];
#pragma unroll
                for (int u = 0; u < 8; ++u) {
                    const V2* xc = xs + (size_t)cur[u] * Bp;
                    if constexpr (TPL == 2) {
                        const float4 q = *reinterpret_cast<const float4*>(xc + 2 * lane);
                        xr[u][0] = q.x; xi[u][0] = q.y; xr[u][1] = q.z; xi[u][1] = q.w;
                    } else {
                        const V2 q = xc[lane];
                        xr[u][0] = q.x; xi[u][0] = q.y;
                    }
                }
                if (p + 16 <= u1) {
#pragma unroll
                    for (int u = 0; u < 8; ++u) nxt[u] = uniform(rc[p + 8 + u]);
                }
                fmag<VT, TPL, kRB, 8>(rv, p, xr, xi, ar, ai);
#pragma unroll
                for (int u = 0; u < 8; ++u) cur[u] = nxt[u];
            }
        }
        for (; p < u1; ++p) {
            const V2* xc = xs + (size_t)uniform(rc[p]) * Bp;
            VT xr[TPL], xi[TPL];
            if constexpr (TPL == 2) {
                const float4 q = *reinterpret_cast<const float4*>(xc + 2 * lane);
                xr[0] = q.x; xi[0] = q.y; xr[1] = q.z; xi[1] = q.w;
            } else {
                const V2 q = xc[lane];
                xr[0] = q.x; xi[0] = q.y;
            }
#pragma unroll
            for (int r = 0; r < kRB; ++r) {
                const double v = rv[(size_t)p * kRB + r];
#pragma unroll
                for (int t = 0; t < TPL; ++t) {
                    ar[t][r] = __builtin_fma(v, (double)xr[t], ar[t][r]);
                    ai[t][r] = __builtin_fma(v, (double)xi[t], ai[t][r]);
                }
            }
        }
        const double g = d.g;
#pragma unroll
        for (int r = 0; r < kRB; ++r) {
            const int i = rb * kRB + r;
            if (i < d.n) {
#pragma unroll
                for (int t = 0; t < TPL; ++t) {
                    const int b = b0 + t;
                    const size_t e = (size_t)i * Bp + b;
                    const V2 xv = x[e];
                    const double di = (double)diag[e];
                    // w = g J x + d x in FP64, rounded once to the vector type
                    VT wr = (VT)__builtin_fma(g, ar[t][r], di * (double)xv.x);
                    VT wi = (VT)__builtin_fma(g, ai[t][r], di * (double)xv.y);
                    if (i >= d.nact[b]) { wr = 0; wi = 0; }
                    if (!last) {
                        V2 w;
                        w.x = wr;
                        w.y = wi;
                        y[e] = w;
                        const double xr = (double)xv.x, xm = (double)xv.y, dr = (double)wr, dm = (double)wi;
                        acc[t][0] = __builtin_fma(xm, dm, __builtin_fma(xr, dr, acc[t][0]));
                        acc[t][1] = __builtin_fma(-xm, dr, __builtin_fma(xr, dm, acc[t][1]));
                    } else {
                        const double dr = (double)wr, dm = (double)wi;
                        acc[t][0] = __builtin_fma(dm, dm, __builtin_fma(dr, dr, acc[t][0]));
                    }
                }
            }
        }
    }
    // partial sums of the workgroup (4 waves = 4 row blocks), per trajectory
    __shared__ double lds[4][2][TPL][64];
#pragma unroll
    for (int t = 0; t < TPL; ++t) { lds[wave][0][t][lane] = acc[t][0]; lds[wave][1][t][lane] = acc[t][1]; }
    __syncthreads();
    if (wave == 0) {
#pragma unroll
        for (int q = 0; q < 2; ++q) {
            if (q == 1 && last) break;
#pragma unroll
            for (int t = 0; t < TPL; ++t) {
                const double v = (lds[0][q][t][lane] + lds[1][q][t][lane]) + (lds[2][q][t][lane] + lds[3][q][t][lane]);
                d.part_s[((size_t)q * d.ntile_s + rtile) * Bp + b0 + t] = v;
            }
        }
    }
}

template <int RB, bool PIPE>
static void launch_spmm_rb(const Dev& d, int j, int last, int tpl, int phases, int first, int ntt, hipStream_t s)
{
    const dim3 grid(bpad8(d.ntile_s * ntt)), block(kThreads);
    for (int ph = 0; ph < phases; ++ph) {
        const int tt0 = first + ph * ntt;
        if (tpl == 2) hipLaunchKernelGGL((k_spmm_b<float, 2, RB, PIPE>), grid, block, 0, s, d, j, last, tt0, ntt);
        else if (d.mixed) hipLaunchKernelGGL((k_spmm_b<float, 1, RB, PIPE>), grid, block, 0, s, d, j, last, tt0, ntt);
        else hipLaunchKernelGGL((k_spmm_b<double, 1, RB, PIPE>), grid, block, 0, s, d, j, last, tt0, ntt);
    }
}

// ---------------------------------------------------------------------------
// Mixed-precision SpMM on the matrix cores.  A block of 16 rows is a dense
// 16 x U matrix over the union of its columns (U padded to a multiple of 16
// with zero columns), so the block's rows of w are that matrix times the U
// gathered rows of x: for a slice of 8 NG trajectories a (16 x U) x (U x 16 NG)
// FP32 product, the 16 NG columns being re/im of the trajectories.  One wave
// per row block and slice; a K-step of 4 union columns is NG
// v_mfma_f32_16x16x4_f32.  Every group of 16 union slots (4 K-steps) adds
// the low parts of its couplings (val - FP32 val, one
// v_mfma_f32_16x16x16bf16 on bfloat16 operands), and the FP32 sums move into
// FP64 ones every two groups, or (EXACT, spmm_exact_sums) after every K-step:
// the FP32 rounding of the couplings (a fixed perturbation of H) does not
// enter w, nor an FP32 running sum over the whole union (whose partial sums
// cancel to the much smaller w); w is rounded to FP32 once.  Over 5000 steps
// of the 12,000-site tube, the FP32 sums of 32 slots leave 1.6e-6 in the
// populations, those of one K-step 1.7e-7, at about a third more SpMM time.  MFMA lanes: A[i][k] from lane i + 16 k, B[k][c]
// from lane c + 16 k, D[4 q + r][c] in register r of lane c + 16 q.  Group g
// holds the real columns NG c + g, so a lane loads NG contiguous floats of one
// x row per step and ends with rows 4 q .. 4 q + 3 of the trajectories
// NG c / 2 .. NG c / 2 + NG / 2 - 1 of its slice.  The loads run kDepth steps
// ahead of the multiplies (column offset, value and x row of a step are
// refilled as soon as the step is consumed).  Each output element is the same
// sequence of MFMA accumulations whatever NG, so the slice width does not
// change a bit of the result.  Against the 4-row kernel: half the gathered
// bytes per row at 2.7x the multiply-adds, which run on the matrix cores.  On
// the 12,000-site dipole tube (B = 256) it takes 43 us (45 us with the FP64
// sums per K-step) against 57-87 us for the row-blocked kernels with their
// FP64 arithmetic (event-timed, whole batch), and inside the captured step it
// leaves the vector ALU to the other slices' updates.
// ---------------------------------------------------------------------------
typedef float mfma_f4 __attribute__((ext_vector_type(4)));
typedef short bf16x4 __attribute__((ext_vector_type(4)));

// two bfloat16 (the high halves of a and b, truncated) in one dword: a low
__device__ __forceinline__ uint32_t bf16_pair(float a, float b)
{
    return __builtin_amdgcn_perm(__float_as_uint(b), __float_as_uint(a), 0x07060302u);
}

// low-part operand of the 16 union slots from slot p (a multiple of 16): the
// four bfloat16 of this lane at byte 512 (p / 16) + 8 lane
__device__ __forceinline__ bf16x4 lo_at(__amdgpu_buffer_rsrc_t r, uint32_t lane_off, int p)
{
    return __builtin_bit_cast(bf16x4, __builtin_amdgcn_raw_buffer_load_b64(r, lane_off, 32 * p, 0));
}

template <int NG> struct XV;
template <> struct XV<2> { using T = float2; };
template <> struct XV<4> { using T = float4; };
template <> struct XV<8> { struct T { float4 a, b; }; };

// x rows through a raw buffer resource (base = the wave's trajectory slice of
// the vector slot, byte offsets below 2^31, reads past num_records return 0).
// Being intrinsics, not plain loads, they stay where they are issued: a plain
// "load here, use next iteration" is turned by the optimiser into a load at
// the top of the next iteration, right before its use.
__device__ __forceinline__ float4 buf_f4(__amdgpu_buffer_rsrc_t r, uint32_t off)
{
    const auto u = __builtin_amdgcn_raw_buffer_load_b128(r, off, 0, 0);
    return make_float4(__uint_as_float(u[0]), __uint_as_float(u[1]), __uint_as_float(u[2]), __uint_as_float(u[3]));
}

// A[i][k] of the step starting at union slot p (uniform byte offset 64 p)
__device__ __forceinline__ float a_at(__amdgpu_buffer_rsrc_t r, uint32_t lane_off, int p)
{
    return __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(r, lane_off, 64 * p, 0));
}

template <int NG>
__device__ __forceinline__ typename XV<NG>::T load_xv(__amdgpu_buffer_rsrc_t r, uint32_t off)
{
    if constexpr (NG == 2) {
        const auto u = __builtin_amdgcn_raw_buffer_load_b64(r, off, 0, 0);
        return make_float2(__uint_as_float(u[0]), __uint_as_float(u[1]));
    } else if constexpr (NG == 4) {
        return buf_f4(r, off);
    } else {
        return typename XV<8>::T{buf_f4(r, off), buf_f4(r, off + 16)};
    }
}

template <int NG>
__device__ __forceinline__ float xv_at(const typename XV<NG>::T& v, int g)
{
    if constexpr (NG == 2) return g == 0 ? v.x : v.y;
    else if constexpr (NG == 4) return g == 0 ? v.x : (g == 1 ? v.y : (g == 2 ? v.z : v.w));
    else {
        const float4& h = g < 4 ? v.a : v.b;
        const int q = g & 3;
        return q == 0 ? h.x : (q == 1 ? h.y : (q == 2 ? h.z : h.w));
    }
}

template <int NG, bool EXACT>
__global__ __launch_bounds__(kThreads) void k_spmm_mfma(Dev d, int j, int last)
{
    constexpr int TW = 8 * NG;                       // trajectories per wave
    constexpr int TL = NG / 2;                       // trajectories per lane
    constexpr int kDepth = 4;
    using XT = typename XV<NG>::T;
    // leading blocks: the Gram-row reduction of update j-1
    const int ntb = gram_task_blocks(d, j, last);
    if ((int)blockIdx.x < ntb) {
        __shared__ double s_red[16][64];
        gram_task<4>(d, j, blockIdx.x, s_red);
        return;
    }
    const int ntt = d.sn / TW;
    const int logical = xcd_tile(blockIdx.x - ntb, gridDim.x - ntb);
    if (logical >= d.ntile_s * ntt) return;
    // row-tile-major: the slices of a row block run on one XCD, so the block's
    // dense values (2.7x the nonzeros) come from HBM once per launch
    const int rtile = logical / ntt, ttile = d.s0 / TW + logical % ntt;
    const int lane = threadIdx.x & 63;
    const int wave = uniform(threadIdx.x >> 6);
    const int kq = lane >> 4, nc = lane & 15;
    const int Bp = d.Bp;
    const float2* __restrict__ x = reinterpret_cast<const float2*>(d.V) + (size_t)j * d.vs;
    float2* __restrict__ y = reinterpret_cast<float2*>(d.V) + (size_t)(j + 1) * d.vs;
    // this wave's trajectory slice of x: row r at byte r * 8 Bp, lane at 4 NG nc
    const float* xs0 = reinterpret_cast<const float*>(x) + (size_t)ttile * 2 * TW;
    const __amdgpu_buffer_rsrc_t xr = __builtin_amdgcn_make_buffer_rsrc(
        (void*)xs0, (short)0, (int)((size_t)d.n * 8 * Bp - (size_t)ttile * 8 * TW), 0x00020000);
    const uint32_t xlane = 4u * NG * nc;
    // dense block values, 16 x 4 floats per step: lane i + 16 k reads A[i][k]
    // at byte 64 p + 4 lane (p = the step's first union slot)
    const __amdgpu_buffer_rsrc_t ar = __builtin_amdgcn_make_buffer_rsrc((void*)d.rb_val32, (short)0, 0x7fffffff,
                                                                        0x00020000);
    const uint32_t alane = 4u * lane;
    const __amdgpu_buffer_rsrc_t lr = __builtin_amdgcn_make_buffer_rsrc((void*)d.rb_lo16, (short)0, 0x7fffffff,
                                                                        0x00020000);
    const uint32_t llane = 8u * lane;
    const int32_t* __restrict__ rc = d.rb_col;
    const uint32_t rowb = 8u * Bp;                   // bytes per row of x
    const int nblk = (d.n + 15) / 16;
    const int rb = rtile * 4 + wave;
    const int b0 = ttile * TW + TL * nc;             // first of this lane's trajectories
    double acc[TL][2];
#pragma unroll
    for (int t = 0; t < TL; ++t) { acc[t][0] = 0.0; acc[t][1] = 0.0; }
    const int slice_nact = uniform(d.nact[ttile * TW]);
    if (rb < nblk && rb * 16 < slice_nact) {
        const int u0 = uniform(d.rb_ptr[rb]), u1 = uniform(d.rb_ptr[rb + 1]);
        // FP32 sums of kSumGroups groups of 16 union slots, then FP64 sums:
        // rows 4 kq + r, real column NG nc + g
        constexpr int kSumGroups = 2;
        int ng = 0;
        mfma_f4 c[NG];
        double s64[NG][4];
#pragma unroll
        for (int g = 0; g < NG; ++g) {
            c[g] = mfma_f4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int r = 0; r < 4; ++r) s64[g][r] = 0.0;
        }
        // column offsets come through the scalar cache, wave-uniform, 4 kDepth
        // (one 64-byte line) per load, issued one loop iteration before they
        // are used: a vector load there would make every gather wait for all
        // the loads in flight before it, and a scalar load used in the step
        // that issues it stalls the wave on every line miss.  The union is
        // padded to a multiple of 4 kDepth columns (and starts on one), so
        // every refill is issued (past the end it reloads a valid group) and
        // the waits on the loads in flight are static.
        static_assert(kDepth == 4, "one int4 of column offsets per step of the group");
        struct Cols { int4 s[kDepth]; };
        auto cols = [&](int p) {
            const int4* q4 = reinterpret_cast<const int4*>(rc + p);
            Cols r;
#pragma unroll
            for (int q = 0; q < kDepth; ++q)
                r.s[q] = make_int4(uniform(q4[q].x), uniform(q4[q].y), uniform(q4[q].z), uniform(q4[q].w));
            return r;
        };
        // byte offsets in 32-bit arithmetic (24-bit multiplies: columns and
        // row strides are far below 2^24)
        auto x_off = [&](int col) { return (uint32_t)__umul24((unsigned)col, rowb) + xlane; };
        // lane group kq takes column kq of a step: selects, no branches
        const bool k_odd = (kq & 1) != 0, k_high = (kq & 2) != 0;
        auto pick = [&](const int4& q4) {
            const int lo = k_odd ? q4.y : q4.x;
            const int hi = k_odd ? q4.w : q4.z;
            return k_high ? hi : lo;
        };
        if (u1 > u0) {
            float av[kDepth];
            XT xv[kDepth];
            bf16x4 alo = lo_at(lr, llane, u0);
            {
                const Cols c0 = cols(u0);
#pragma unroll
                for (int q = 0; q < kDepth; ++q) {
                    av[q] = a_at(ar, alane, u0 + 4 * q);
                    xv[q] = load_xv<NG>(xr, x_off(pick(c0.s[q])));
                    __builtin_amdgcn_sched_barrier(0);
                }
            }
            // refill group of this iteration (columns pnext .. pnext + 4 kDepth - 1)
            int pnext = u0 + 4 * kDepth < u1 ? u0 + 4 * kDepth : u0;
            Cols cur = cols(pnext);
            for (int base = u0; base < u1; base += 4 * kDepth) {
                const int pn2 = base + 8 * kDepth < u1 ? base + 8 * kDepth : u0;
                const Cols nxt = cols(pn2);
                const bf16x4 alo_n = lo_at(lr, llane, pnext);
                // the group's x values as the bfloat16 operand of the low parts
                // (lane group kq: the slot it multiplied in step q; truncated,
                // which costs the low parts, themselves below 2^-24 of a
                // coupling, 2^-8 of their size)
                float xq[NG][kDepth];
#pragma unroll
                for (int q = 0; q < kDepth; ++q) {
#pragma unroll
                    for (int g = 0; g < NG; ++g) {
                        c[g] = __builtin_amdgcn_mfma_f32_16x16x4f32(av[q], xv_at<NG>(xv[q], g), c[g], 0, 0, 0);
                        xq[g][q] = xv_at<NG>(xv[q], g);
                    }
                    if constexpr (EXACT) {
                        // each K-step's FP32 sum of 4 products into the FP64 sums
#pragma unroll
                        for (int g = 0; g < NG; ++g) {
#pragma unroll
                            for (int r = 0; r < 4; ++r) s64[g][r] += (double)c[g][r];
                            c[g] = mfma_f4{0.f, 0.f, 0.f, 0.f};
                        }
                    }
                    av[q] = a_at(ar, alane, pnext + 4 * q);
                    xv[q] = load_xv<NG>(xr, x_off(pick(cur.s[q])));
                    // step order kept: the refill of step q issues before the
                    // multiplies of step q + 1, so kDepth steps stay in flight
                    __builtin_amdgcn_sched_barrier(0);
                }
#pragma unroll
                for (int g = 0; g < NG; ++g) {
                    const uint32_t lo2 = bf16_pair(xq[g][0], xq[g][1]), hi2 = bf16_pair(xq[g][2], xq[g][3]);
                    const bf16x4 xb = __builtin_bit_cast(bf16x4, (uint2){lo2, hi2});
                    c[g] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(alo, xb, c[g], 0, 0, 0);
                }
                // every kSumGroups groups (and after the last) the FP32 sums
                // move into the FP64 ones
                if (EXACT || ++ng == kSumGroups || base + 4 * kDepth >= u1) {
                    ng = 0;
#pragma unroll
                    for (int g = 0; g < NG; ++g) {
#pragma unroll
                        for (int r = 0; r < 4; ++r) s64[g][r] += (double)c[g][r];
                        c[g] = mfma_f4{0.f, 0.f, 0.f, 0.f};
                    }
                }
                alo = alo_n;
                pnext = pn2;
                cur = nxt;
            }
        }
        const double g = d.g;
        const float* __restrict__ diag = reinterpret_cast<const float*>(d.diag);
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int r = rb * 16 + 4 * kq + i;
            if (r < d.n) {
                const size_t e = (size_t)r * Bp + b0;
                float xs[2 * TL], dv[TL], w[2 * TL];
#pragma unroll
                for (int t = 0; t < TL; ++t) {
                    const float2 q = x[e + t];
                    xs[2 * t] = q.x;
                    xs[2 * t + 1] = q.y;
                    dv[t] = diag[e + t];
                }
#pragma unroll
                for (int t = 0; t < TL; ++t) {
                    // w = g J x + d x in FP64, rounded once to FP32
                    float wr = (float)__builtin_fma(g, s64[2 * t][i], (double)dv[t] * (double)xs[2 * t]);
                    float wi = (float)__builtin_fma(g, s64[2 * t + 1][i], (double)dv[t] * (double)xs[2 * t + 1]);
                    if (r >= d.nact[b0 + t]) { wr = 0.f; wi = 0.f; }
                    w[2 * t] = wr;
                    w[2 * t + 1] = wi;
                    const double xr = (double)xs[2 * t], xm = (double)xs[2 * t + 1], dr = (double)wr, dm = (double)wi;
                    if (!last) {
                        acc[t][0] = __builtin_fma(xm, dm, __builtin_fma(xr, dr, acc[t][0]));
                        acc[t][1] = __builtin_fma(-xm, dr, __builtin_fma(xr, dm, acc[t][1]));
                    } else {
                        acc[t][0] = __builtin_fma(dm, dm, __builtin_fma(dr, dr, acc[t][0]));
                    }
                }
                if (!last) {
#pragma unroll
                    for (int t = 0; t < TL; ++t) y[e + t] = make_float2(w[2 * t], w[2 * t + 1]);
                }
            }
        }
    }
    // rows of the block: the four lane groups kq, then the four waves
#pragma unroll
    for (int t = 0; t < TL; ++t)
#pragma unroll
        for (int q = 0; q < 2; ++q) {
            acc[t][q] += __shfl_xor(acc[t][q], 16, 64);
            acc[t][q] += __shfl_xor(acc[t][q], 32, 64);
        }
    __shared__ double lds[4][2][TL][16];
    if (kq == 0) {
#pragma unroll
        for (int t = 0; t < TL; ++t) { lds[wave][0][t][nc] = acc[t][0]; lds[wave][1][t][nc] = acc[t][1]; }
    }
    __syncthreads();
    if (wave == 0 && kq == 0) {
#pragma unroll
        for (int q = 0; q < 2; ++q) {
            if (q == 1 && last) break;
#pragma unroll
            for (int t = 0; t < TL; ++t) {
                const double v = (lds[0][q][t][nc] + lds[1][q][t][nc]) + (lds[2][q][t][nc] + lds[3][q][t][nc]);
                d.part_s[((size_t)q * d.ntile_s + rtile) * Bp + b0 + t] = v;
            }
        }
    }
}

hipError_t launch_spmm_b(const Dev& d, int j, int last, hipStream_t s)
{
    if (d.ring_rows > 0) return launch_spmm_ring(d, j, last, s);
    if (d.rb == 16) {
        // the matrix-core kernel (mixed precision only, enforced by the tuning
        // checks); spmm_traj_per_lane 1, 2, 4 = 16, 32, 64 trajectories per wave
        const int tl = d.spmm_tpl;
        const int tw = 16 * tl;
        const dim3 grid(gram_task_blocks(d, j, last) + bpad8(d.ntile_s * (d.sn / tw))), block(kThreads);
        if (d.spmm_exact) {
            if (tl == 4) hipLaunchKernelGGL((k_spmm_mfma<8, true>), grid, block, 0, s, d, j, last);
            else if (tl == 2) hipLaunchKernelGGL((k_spmm_mfma<4, true>), grid, block, 0, s, d, j, last);
            else hipLaunchKernelGGL((k_spmm_mfma<2, true>), grid, block, 0, s, d, j, last);
        } else {
            if (tl == 4) hipLaunchKernelGGL((k_spmm_mfma<8, false>), grid, block, 0, s, d, j, last);
            else if (tl == 2) hipLaunchKernelGGL((k_spmm_mfma<4, false>), grid, block, 0, s, d, j, last);
            else hipLaunchKernelGGL((k_spmm_mfma<2, false>), grid, block, 0, s, d, j, last);
        }
        return hipGetLastError();
    }
    // trajectory slices of 64 * TPL; the slices run as `phases` launches of
    // ntt slices each, so that one launch's gathered rows fit the L2 of an XCD
    const int tpl = (d.mixed && d.spmm_tpl == 2 && d.sn % 128 == 0) ? 2 : 1;
    const int nslice = d.sn / (64 * tpl), first = d.s0 / (64 * tpl);
    int phases = d.spmm_phases < 1 ? 1 : d.spmm_phases;
    while (nslice % phases) --phases;
    const int ntt = nslice / phases;
    const bool pipe = d.spmm_pipe != 0;
    switch (d.rb) {
    case 2: pipe ? launch_spmm_rb<2, true>(d, j, last, tpl, phases, first, ntt, s) : launch_spmm_rb<2, false>(d, j, last, tpl, phases, first, ntt, s); break;
    case 8: pipe ? launch_spmm_rb<8, true>(d, j, last, tpl, phases, first, ntt, s) : launch_spmm_rb<8, false>(d, j, last, tpl, phases, first, ntt, s); break;
    default: pipe ? launch_spmm_rb<4, true>(d, j, last, tpl, phases, first, ntt, s) : launch_spmm_rb<4, false>(d, j, last, tpl, phases, first, ntt, s); break;
    }
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Taylor part of a Taylor-Krylov cycle (taylorKrylovClass.cu:226-322), batched:
// u_k = A u_{k-1} with the row-blocked couplings of the SpMM above, FP64
// vectors (|A^k psi| grows as |A|^k), u_0 = psi; t = psi + c u_1 at k = 1,
// t += c^k u_k / k! for 1 < k < p; at k = p the partial |u_p|^2 per SpMM tile.
// ---------------------------------------------------------------------------
template <int kRB>
__global__ __launch_bounds__(kThreads) void k_tk_power_rb(Dev d, int k)
{
    const int ntt = d.sn / 64;
    const int logical = xcd_tile(blockIdx.x, gridDim.x);
    if (logical >= d.ntile_s * ntt) return;
    const int rtile = logical / ntt, ttile = d.s0 / 64 + logical % ntt;
    const int lane = threadIdx.x & 63;
    const int wave = uniform(threadIdx.x >> 6);
    const int Bp = d.Bp;
    const int b = ttile * 64 + lane;
    const size_t nB = (size_t)d.n * Bp;
    const double2* __restrict__ x = k == 1 ? d.psi : d.tk_u + (size_t)((k - 1) & 1) * nB;
    double2* __restrict__ y = d.tk_u + (size_t)(k & 1) * nB;
    const double2* __restrict__ xs = x + (size_t)ttile * 64;
    const double* __restrict__ rv = d.rb_val64;
    const int32_t* __restrict__ rc = d.rb_col;
    const bool last = k == d.tk_p;
    const double2 ck = last ? make_double2(0.0, 0.0) : d.tk_coef[k];
    const int nblk = (d.n + kRB - 1) / kRB;
    const int rb = rtile * 4 + wave;
    double acc = 0.0;
    const int slice_nact = uniform(d.nact[ttile * 64]);
    if (rb < nblk && rb * kRB < slice_nact) {
        const int u0 = uniform(d.rb_ptr[rb]), u1 = uniform(d.rb_ptr[rb + 1]);
        double ar[1][kRB], ai[1][kRB];
#pragma unroll
        for (int r = 0; r < kRB; ++r) { ar[0][r] = 0.0; ai[0][r] = 0.0; }
        int p = u0;
        for (; p + 8 <= u1; p += 8) {
            double xr[8][1], xi[8][1];
            gatherg<double, 1, 8>(xs, rc, p, Bp, lane, xr, xi);
            fmag<double, 1, kRB, 8>(rv, p, xr, xi, ar, ai);
        }
        for (; p < u1; ++p) {
            const double2 q = xs[(size_t)uniform(rc[p]) * Bp + lane];
#pragma unroll
            for (int r = 0; r < kRB; ++r) {
                const double v = rv[(size_t)p * kRB + r];
                ar[0][r] = fmav(v, q.x, ar[0][r]);
                ai[0][r] = fmav(v, q.y, ai[0][r]);
            }
        }
        const int nact = d.nact[b];
#pragma unroll
        for (int r = 0; r < kRB; ++r) {
            const int i = rb * kRB + r;
            if (i < d.n) {
                const size_t e = (size_t)i * Bp + b;
                const double2 xv = x[e];
                const double di = d.mixed ? (double)reinterpret_cast<const float*>(d.diag)[e]
                                          : reinterpret_cast<const double*>(d.diag)[e];
                double wr = fmav(d.g, ar[0][r], di * xv.x);
                double wi = fmav(d.g, ai[0][r], di * xv.y);
                if (i >= nact) { wr = 0.0; wi = 0.0; }
                y[e] = make_double2(wr, wi);
                if (!last) {
                    double2 tv = k == 1 ? xv : d.tk_t[e];
                    tv.x += ck.x * wr - ck.y * wi;
                    tv.y += ck.x * wi + ck.y * wr;
                    d.tk_t[e] = tv;
                } else {
                    acc += wr * wr + wi * wi;
                }
            }
        }
    }
    if (last) {
        __shared__ double lds[4][64];
        lds[wave][lane] = acc;
        __syncthreads();
        if (wave == 0)
            d.part[(size_t)rtile * Bp + b] = (lds[0][lane] + lds[1][lane]) + (lds[2][lane] + lds[3][lane]);
    }
}

hipError_t launch_tk_power_b(const Dev& d, int k, hipStream_t s)
{
    const dim3 grid(bpad8(d.ntile_s * (d.sn / 64))), block(kThreads);
    switch (d.rb) {
    case 2: hipLaunchKernelGGL((k_tk_power_rb<2>), grid, block, 0, s, d, k); break;
    case 8: hipLaunchKernelGGL((k_tk_power_rb<8>), grid, block, 0, s, d, k); break;
    case 16: hipLaunchKernelGGL((k_tk_power_rb<16>), grid, block, 0, s, d, k); break;
    default: hipLaunchKernelGGL((k_tk_power_rb<4>), grid, block, 0, s, d, k); break;
    }
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Per-trajectory scalars of iteration j (batched): grid (Bp/64, ceil(j/NW) or
// 1), NW waves.  Block y == 0 takes <v_j|w> -- FUSED (after the matrix-core
// SpMM, NW = 8): summed over the SpMM tiles by the block itself, both halves
// at once; else from red[kRedC..] (k_reduce_all, NW = 4), the same order of
// tile_sum16 either way -- and stores alpha_j, beta_{j-1}, s_j and the Gram
// row of its 64 trajectories, rebuilt from the reduced Gram sums of update
// j-1.  Every block rebuilds s_j and the Gram row for itself (wave w the
// entries l in [8w, 8w + 8)); wave w of block y then computes the projection
// h_k, k = NW y + w < j, which needs no <v_j|w>.  The 512-thread fused block
// (not 1024) finds room on CUs busy with the other slices' updates sooner.
// ---------------------------------------------------------------------------
template <bool FUSED>
__global__ __launch_bounds__(FUSED ? 512 : 256) void k_scalars_hb(Dev d, int j)
{
    constexpr int NW = FUSED ? 8 : 4;
    __shared__ double2 s_G[kMaxM + 1][64];
    __shared__ double s_red[FUSED ? 32 : 1][64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int b0 = d.s0 + blockIdx.x * 64, b = b0 + lane;
    const int Bp = d.Bp;
    double cr = 0.0, ci = 0.0;
    if constexpr (FUSED) {
        if (blockIdx.y == 0) {
            const int maxn = min(d.n, uniform(d.nact[b0]));
            const int th = d.ring_rows > 0 ? d.ring_rows : d.rt_s;
            const int nts = min(d.ntile_p, (maxn + th - 1) / th);
            tile_sum16_pair<NW>(d.part_s + b, d.part_s + (size_t)d.ntile_p * Bp + b, nts, Bp, s_red, cr, ci);
        }
    } else {
        cr = d.red[((size_t)kRedC(d.m)) * Bp + b];
        ci = d.red[((size_t)kRedC(d.m) + 1) * Bp + b];
    }
    const bool store = blockIdx.y == 0 && wave == 0;
    // s_j = 1 / beta_{j-1}: every wave from the same loads (broke only ever
    // goes 0 -> 1 on this very test, so a wave that reads the flag after wave
    // 0 stored it reaches the same result)
    double sj = 1.0, bet = 0.0;
    if (j == 0) {
        if (store) { d.sc[b] = 1.0; d.broke[b] = 0; }
    } else {
        const double nu = d.red[(size_t)(2 * j) * Bp + b];
        bet = __builtin_sqrt(nu);
        const double bprev = j >= 2 ? d.beta[(size_t)(j - 2) * Bp + b] : 0.0;
        const double aprev = d.alpha[(size_t)(j - 1) * Bp + b];
        int broke = d.broke[b];
        if (broke || !(bet > d.tol * (__builtin_fabs(aprev) + bprev))) {
            broke = 1;
            bet = 0.0;
            sj = 0.0;
        } else {
            sj = 1.0 / bet;
        }
        if (store) {
            d.broke[b] = broke;
            d.beta[(size_t)(j - 1) * Bp + b] = bet;
            d.sc[(size_t)j * Bp + b] = sj;
        }
    }
    if (store) {
        const double hr = sj * sj * cr, hi = sj * sj * ci;
        d.alpha[(size_t)j * Bp + b] = hr;
        d.hco[((size_t)j * kMaxM + j) * Bp + b] = make_double2(hr, hi);
    }
    // Gram row of v_j rescaled: wave w takes l in [8 w, 8 w + 8) (+ 8 NW r),
    // the eight loads of a group issued together
    for (int l0 = 8 * wave; l0 < j; l0 += 8 * NW) {
        double sl[8], gr[8], gi[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const int l = min(l0 + u, j - 1);
            sl[u] = d.sc[(size_t)l * Bp + b];
            gr[u] = d.red[(size_t)(2 * l) * Bp + b];
            gi[u] = d.red[(size_t)(2 * l + 1) * Bp + b];
        }
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const int l = l0 + u;
            if (l < j) {
                const double f = sl[u] * sj;
                const double2 G = make_double2(f * gr[u], f * gi[u]);
                s_G[l][lane] = G;
                if (blockIdx.y == 0) d.gram[(size_t)l * Bp + b] = G;
            }
        }
    }
    if (wave == 0) s_G[j][lane] = make_double2(1.0, 0.0);
    __syncthreads();
    const int k = blockIdx.y * NW + wave;
    if (k < j) {
        const double bk = (k == j - 1) ? bet : d.beta[(size_t)k * Bp + b];
        const double2 Gk1 = s_G[k + 1][lane];
        double sr = bk * Gk1.x, si = bk * Gk1.y;
        // h^(k)_l in rounds of eight independent loads, summed in order of l
        for (int l0 = 0; l0 <= k; l0 += 8) {
            double2 a[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) a[u] = d.hco[((size_t)k * kMaxM + min(l0 + u, k)) * Bp + b];
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                const int l = l0 + u;
                if (l <= k) {
                    const double2 G = s_G[l][lane];
                    sr += a[u].x * G.x + a[u].y * G.y;
                    si += a[u].x * G.y - a[u].y * G.x;
                }
            }
        }
        d.hco[((size_t)j * kMaxM + k) * Bp + b] = make_double2(sr, si);
    }
}

// out[b] = sum over tiles t < ntile of p[t Bp + b] (p = one quantity's tile
// column): 16 waves per block, eight independent loads in flight per lane,
// fixed summation order; block = 64 trajectories starting at b - lane
__device__ __forceinline__ void wide_tile_sum(const double* p, int ntile, int Bp, double* out)
{
    __shared__ double lds[16][64];
    const double s = tile_sum16<16>(p, ntile, Bp, lds);
    if ((threadIdx.x >> 6) == 0) *out = s;
}

// red[dst + q][b] = sum over tiles of part[q][tile][b]
__global__ __launch_bounds__(1024) void k_reduce_wide(Dev d, int ntile, int dst)
{
    const int q = blockIdx.y;
    const int b = d.s0 + blockIdx.x * 64 + (threadIdx.x & 63);
    const int Bp = d.Bp;
    wide_tile_sum(d.part + (size_t)q * ntile * Bp + b, ntile, Bp, d.red + ((size_t)dst + q) * Bp + b);
}

// after the SpMMs that do not carry the Gram reduction: blockIdx.y < 2:
// <v_j|w> (re, im) from the SpMM tiles into red[kRedC..]; else quantity y - 2
// of update j-1 (Gram row re/im pairs, then |v_j|^2) for one group of 64
// slots, as the matrix-core SpMM's leading blocks do it (gram_task)
__global__ __launch_bounds__(1024) void k_reduce_all(Dev d, int j)
{
    __shared__ double lds[16][64];
    const int y = blockIdx.y;
    if (y >= 2) {
        gram_task<16>(d, j, blockIdx.x * (2 * j + 1) + (y - 2), lds);
        return;
    }
    const int b0 = d.s0 + blockIdx.x * 64, b = b0 + (threadIdx.x & 63);
    const int maxn = min(d.n, uniform(d.nact[b0]));
    const int th = d.ring_rows > 0 ? d.ring_rows : d.rt_s;
    const double v = tile_sum16<16>(d.part_s + (size_t)y * d.ntile_p * d.Bp + b, min(d.ntile_p, (maxn + th - 1) / th),
                                    d.Bp, lds);
    if ((threadIdx.x >> 6) == 0) d.red[((size_t)kRedC(d.m) + y) * d.Bp + b] = v;
}

hipError_t launch_reduce_all(const Dev& d, int j, hipStream_t s)
{
    const int nq = 2 + (j >= 1 ? 2 * j + 1 : 0);
    hipLaunchKernelGGL(k_reduce_all, dim3(d.sn / 64, nq), dim3(1024), 0, s, d, j);
    return hipGetLastError();
}

hipError_t launch_reduce_wide(const Dev& d, int nq, int ntile, int dst, hipStream_t s)
{
    hipLaunchKernelGGL(k_reduce_wide, dim3(d.sn / 64, nq), dim3(1024), 0, s, d, ntile, dst);
    return hipGetLastError();
}

// after the SpMM of iteration j (and launch_reduce_all when it does not carry
// the Gram reduction)
hipError_t launch_scalars_hb(const Dev& d, int j, hipStream_t s)
{
    if (spmm_fuses_gram(d)) {
        const int ny = j > 0 ? (j + 7) / 8 : 1;
        hipLaunchKernelGGL(k_scalars_hb<true>, dim3(d.sn / 64, ny), dim3(512), 0, s, d, j);
    } else {
        const int ny = j > 0 ? (j + 3) / 4 : 1;
        hipLaunchKernelGGL(k_scalars_hb<false>, dim3(d.sn / 64, ny), dim3(kThreads), 0, s, d, j);
    }
    return hipGetLastError();
}

// resident waves per SIMD asked of the register allocator for the FP32
// instantiations (each fits without spilling; one more wave than the
// allocator's own choice for KW = 3, 4, 5, 6)
template <typename VT, int KW>
constexpr int kUpdateWaves = sizeof(VT) != 4 ? 1 : (KW <= 2 ? 8 : KW == 3 ? 6 : KW <= 5 ? 5 : KW == 6 ? 4 : 1);

// a basis element read with the nontemporal hint
template <typename T2>
__device__ __forceinline__ T2 load_nt(const T2* p)
{
    T2 r;
    r.x = __builtin_nontemporal_load(&p->x);
    r.y = __builtin_nontemporal_load(&p->y);
    return r;
}

// ---------------------------------------------------------------------------
// Update of iteration j: v_{j+1} = s_j w - sum_{k<=j} (h_k s_k) v_k and the
// FP64 partial Gram row <v_k|v_{j+1}>, |v_{j+1}|^2.  Block = rt_e rows x 16
// trajectories; lane = 16 kg + t handles vectors k = kg, kg+4, ... (KW =
// ceil((j+1)/4) of them, so only the last group can hold a k > j).  A wave
// takes rows r0 + wave, + 4, + 8, ...: two row buffers, each refilled 8 rows
// ahead right after it is consumed, so the loads of two rows are in flight
// while one is combined (a single buffer copied forward each row would make
// every row wait for the loads just issued).  The coefficients -h_k s_k are
// staged once per block in LDS.
//
// The NT oldest vector groups u < NT (vectors 4u .. 4u+3) are read with the
// nontemporal hint, the others with cached loads: the launcher keeps the
// newest d.upd_cached groups cached, which fit in the last-level cache
// (kUpdCacheBytes) and are re-read first, and streaming the older ones with
// the hint keeps them from evicting the newest.  With every vector hinted, or
// none, the step is 7-10 % slower at N = 12,000, B = 256.
//
// The new rows are not stored as they are made: a store counts in the same
// wait counter as the loads, so each row's wait would also wait for the
// previous row's write.  Each wave collects its rows in LDS and stores 16 of
// them at a time (4 rows per store instruction); that LDS is the same that
// later combines the four waves' Gram sums.
// ---------------------------------------------------------------------------
template <typename VT, int KW, int NT>
__global__ __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(kUpdateWaves<VT, KW>, 8)))
void k_update_b(Dev d, int j)
{
    using V2 = typename BVec2<VT>::T;
    constexpr int kStage = 16;                                  // rows collected per wave
    constexpr int kRedD = (2 * KW + 1) * 64;                    // doubles: Gram sums of a wave
    constexpr int kStageD = kStage * 16 * (int)sizeof(V2) / 8;  // doubles: collected rows
    __shared__ double2 s_c[4 * KW][16];
    __shared__ double s_w[16];
    __shared__ double wl[4][kRedD > kStageD ? kRedD : kStageD];
    const int ntt = d.sn / 16;
    const int logical = xcd_tile(blockIdx.x, gridDim.x);
    if (logical >= d.ntile_e * ntt) return;
    const int rtile = logical / ntt, ttile = d.s0 / 16 + logical % ntt;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int t = lane & 15, kg = lane >> 4;
    const int b = ttile * 16 + t;
    const int Bp = d.Bp;
    const V2* __restrict__ V = reinterpret_cast<const V2*>(d.V);
    V2* __restrict__ out = reinterpret_cast<V2*>(d.V) + (size_t)(j + 1) * d.vs;
    // rows at or beyond the first (largest) n_active of the 16 sorted slots
    // are inactive for all of them: skipped, they stay 0
    const int r0 = rtile * d.rt_e, r1 = min(min(d.n, r0 + d.rt_e), (int)uniform(d.nact[ttile * 16]));
    // every load is issued unconditionally (row index clamped to the wave's
    // last row, vector index to j, w read by all four lane groups), so the
    // wait before a row is "all but the other buffer's loads", never a drain
    const int i0 = r0 + wave;
    const int nrw = r1 > i0 ? (r1 - i0 + 3) / 4 : 0;       // rows of this wave
    auto load_row = [&](int n, V2 (&v)[KW], V2& w) {
        const size_t e = (size_t)(i0 + 4 * min(n, nrw - 1)) * Bp + b;
#pragma unroll
        for (int u = 0; u < NT; ++u) v[u] = load_nt(&V[(size_t)min(kg + 4 * u, j) * d.vs + e]);
#pragma unroll
        for (int u = NT; u < KW; ++u) v[u] = V[(size_t)min(kg + 4 * u, j) * d.vs + e];
        w = out[e];
    };
    // the first two rows' loads are issued before the coefficients are staged
    V2 va[KW], vb[KW], wa, wb;
    if (nrw > 0) {
        load_row(0, va, wa);
        load_row(1, vb, wb);
    }
    for (int q = threadIdx.x; q < 16 * (j + 1); q += kThreads) {
        const int k = q >> 4, tt = q & 15, bb = ttile * 16 + tt;
        const double2 h = d.hco[((size_t)j * kMaxM + k) * Bp + bb];
        const double sk = d.sc[(size_t)k * Bp + bb];
        s_c[k][tt] = make_double2(-h.x * sk, -h.y * sk);
        if (k == 0) s_w[tt] = d.sc[(size_t)j * Bp + bb];
    }
    __syncthreads();
    const double sw = s_w[t];
    double acc[2 * KW + 1];
#pragma unroll
    for (int q = 0; q < 2 * KW + 1; ++q) acc[q] = 0.0;
    V2* st = reinterpret_cast<V2*>(&wl[wave][0]);               // [kStage][16] rows of this wave
    auto do_row = [&](int n, const V2 (&vk)[KW], const V2& w) {
        // for KW >= 5 the coefficients are read from LDS for every row
        // rather than held in registers across the loop (no spills at the
        // requested occupancy)
        if (KW >= 5) asm volatile("" ::: "memory");
        double pr = 0.0, pi = 0.0;
        if (kg == 0) {
            pr = sw * (double)w.x;
            pi = sw * (double)w.y;
        }
#pragma unroll
        for (int u = 0; u < KW; ++u) {
            const int k = kg + 4 * u;
            if (u < KW - 1 || k <= j) {
                const double2 cf = s_c[k][t];
                const double xr = (double)vk[u].x, xi = (double)vk[u].y;
                pr += cf.x * xr - cf.y * xi;
                pi += cf.x * xi + cf.y * xr;
            }
        }
        pr += __shfl_xor(pr, 16, 64);
        pi += __shfl_xor(pi, 16, 64);
        pr += __shfl_xor(pr, 32, 64);
        pi += __shfl_xor(pi, 32, 64);
        V2 nv;
        nv.x = (VT)pr;
        nv.y = (VT)pi;
        if (kg == 0) st[(n % kStage) * 16 + t] = nv;
        const double nr = (double)nv.x, ni = (double)nv.y;
#pragma unroll
        for (int u = 0; u < KW; ++u) {
            const int k = kg + 4 * u;
            if (u < KW - 1 || k <= j) {
                const double xr = (double)vk[u].x, xi = (double)vk[u].y;
                acc[2 * u] += xr * nr + xi * ni;
                acc[2 * u + 1] += xr * ni - xi * nr;
            }
        }
        if (kg == 0) acc[2 * KW] += nr * nr + ni * ni;
    };
    // rows n0 .. n0 + kStage - 1 (those below nrw) from LDS to the vector:
    // lane group kg stores rows n0 + kg, + 4, + 8, + 12
    auto flush = [&](int n0) {
#pragma unroll
        for (int s2 = 0; s2 < kStage / 4; ++s2) {
            const int n = n0 + kg + 4 * s2;
            if (n < nrw) out[(size_t)(i0 + 4 * n) * Bp + b] = st[(n % kStage) * 16 + t];
        }
    };
    for (int n = 0; n < nrw; n += 2) {
        do_row(n, va, wa);
        load_row(n + 2, va, wa);
        if (n + 1 < nrw) do_row(n + 1, vb, wb);
        load_row(n + 3, vb, wb);
        if (n % kStage == kStage - 2 || n + 2 >= nrw) flush(n - n % kStage);
    }
    // combine the four waves (rows) per lane, then write q = 2k, 2k+1 and nu
    // (each wave overwrites only its own LDS, after its own last flush read it)
#pragma unroll
    for (int q = 0; q < 2 * KW + 1; ++q) wl[wave][q * 64 + lane] = acc[q];
    __syncthreads();
    if (wave == 0) {
        auto sum4 = [&](int q) {
            return (wl[0][q * 64 + lane] + wl[1][q * 64 + lane]) + (wl[2][q * 64 + lane] + wl[3][q * 64 + lane]);
        };
#pragma unroll
        for (int u = 0; u < KW; ++u) {
            const int k = kg + 4 * u;
            if (k <= j) {
                for (int c = 0; c < 2; ++c)
                    d.part[((size_t)(2 * k + c) * d.ntile_e + rtile) * Bp + b] = sum4(2 * u + c);
            }
        }
        if (kg == 0) d.part[((size_t)(2 * (j + 1)) * d.ntile_e + rtile) * Bp + b] = sum4(2 * KW);
    }
}

// k_update_b<VT, KW, nt> for the run-time nt in [0, KW]
template <typename VT, int KW, int NT = KW>
static void launch_update_nt(const Dev& d, int j, int nt, dim3 grid, hipStream_t s)
{
    if constexpr (NT > 0) {
        if (nt < NT) {
            launch_update_nt<VT, KW, NT - 1>(d, j, nt, grid, s);
            return;
        }
    }
    if (!launch_with_events(k_update_b<VT, KW, NT>, grid, dim3(kThreads), s, d, j))
        hipLaunchKernelGGL((k_update_b<VT, KW, NT>), grid, dim3(kThreads), 0, s, d, j);
}

template <typename VT, int KW>
static void launch_update_kw(const Dev& d, int j, dim3 grid, hipStream_t s)
{
    launch_update_nt<VT, KW>(d, j, std::max(0, KW - d.upd_cached), grid, s);
}

template <typename VT>
static void launch_update_bt(const Dev& d, int j, hipStream_t s)
{
    const dim3 grid(bpad8(d.ntile_e * (d.sn / 16)));
    switch ((j + 4) / 4) {          // KW = ceil((j + 1) / 4)
    case 1: launch_update_kw<VT, 1>(d, j, grid, s); break;
    case 2: launch_update_kw<VT, 2>(d, j, grid, s); break;
    case 3: launch_update_kw<VT, 3>(d, j, grid, s); break;
    case 4: launch_update_kw<VT, 4>(d, j, grid, s); break;
    case 5: launch_update_kw<VT, 5>(d, j, grid, s); break;
    case 6: launch_update_kw<VT, 6>(d, j, grid, s); break;
    case 7: launch_update_kw<VT, 7>(d, j, grid, s); break;
    default: launch_update_kw<VT, 8>(d, j, grid, s); break;
    }
}

hipError_t launch_update_b(const Dev& d, int j, hipStream_t s)
{
    if (d.mixed) launch_update_bt<float>(d, j, s);
    else launch_update_bt<double>(d, j, s);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Combine psi = y_0 psi + sum_{k>=1} (y_k s_k) v_k (propStep.cu:31-44) and the
// partial |psi|^2; same 16 x 4 lane layout as the update.  TK: the Krylov part
// of a Taylor-Krylov cycle, psi = t + sum_{k>=0} (y_k s_k) v_k
// (taylorKrylovClass.cu:372-418).  The basis is read with the nontemporal
// hint: no kernel reads it again before the next cycle rewrites it, and the
// hint keeps it from evicting psi, which the next prologue reads.
// ---------------------------------------------------------------------------
template <typename VT, int KW, bool TK>
__global__ __launch_bounds__(kThreads) void k_combine_b(Dev d)
{
    constexpr int k0 = TK ? 0 : 1;
    using V2 = typename BVec2<VT>::T;
    __shared__ double2 s_y[4 * KW][16];
    __shared__ double lds[4][64];
    const int ntt = d.sn / 16;
    const int logical = xcd_tile(blockIdx.x, gridDim.x);
    if (logical >= d.ntile_e * ntt) return;
    const int rtile = logical / ntt, ttile = d.s0 / 16 + logical % ntt;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int t = lane & 15, kg = lane >> 4;
    const int b = ttile * 16 + t;
    const int Bp = d.Bp, m = d.m;
    const size_t nB = (size_t)d.n * Bp;
    for (int q = threadIdx.x; q < 16 * m; q += kThreads) {
        const int k = q >> 4, tt = q & 15;
        s_y[k][tt] = d.ycoef[(size_t)k * Bp + ttile * 16 + tt];
    }
    __syncthreads();
    double2 y[KW];
#pragma unroll
    for (int u = 0; u < KW; ++u) {
        const int k = kg + 4 * u;
        y[u] = (k < m) ? s_y[k][t] : make_double2(0.0, 0.0);
    }
    const V2* __restrict__ V = reinterpret_cast<const V2*>(d.V);
    double acc = 0.0;
    // rows at or beyond the first (largest) n_active of the 16 sorted slots
    // are inactive for all of them: skipped, they stay 0
    const int r0 = rtile * d.rt_e, r1 = min(min(d.n, r0 + d.rt_e), (int)uniform(d.nact[ttile * 16]));
    V2 nk[KW];
    double2 np = make_double2(0.0, 0.0);
    if (r0 + wave < r1) {
        const size_t e = (size_t)(r0 + wave) * Bp + b;
#pragma unroll
        for (int u = 0; u < KW; ++u) {
            const int k = kg + 4 * u;
            if (k >= k0 && k < m) nk[u] = load_nt(&V[(size_t)k * d.vs + e]);
        }
        if (kg == 0) np = TK ? d.tk_t[e] : d.psi[e];
    }
    for (int i = r0 + wave; i < r1; i += 4) {
        const size_t e = (size_t)i * Bp + b;
        V2 vk[KW];
#pragma unroll
        for (int u = 0; u < KW; ++u) vk[u] = nk[u];
        const double2 p = np;
        if (i + 4 < r1) {
            const size_t en = e + (size_t)4 * Bp;
#pragma unroll
            for (int u = 0; u < KW; ++u) {
                const int k = kg + 4 * u;
                if (k >= k0 && k < m) nk[u] = load_nt(&V[(size_t)k * d.vs + en]);
            }
            if (kg == 0) np = TK ? d.tk_t[en] : d.psi[en];
        }
        double pr = 0.0, pi = 0.0;
        if (kg == 0) {
            if constexpr (TK) {
                pr = p.x;
                pi = p.y;
            } else {
                pr = y[0].x * p.x - y[0].y * p.y;
                pi = y[0].x * p.y + y[0].y * p.x;
            }
        }
#pragma unroll
        for (int u = 0; u < KW; ++u) {
            const int k = kg + 4 * u;
            if (k >= k0 && k < m) {
                const double xr = (double)vk[u].x, xi = (double)vk[u].y;
                pr += y[u].x * xr - y[u].y * xi;
                pi += y[u].x * xi + y[u].y * xr;
            }
        }
        pr += __shfl_xor(pr, 16, 64);
        pi += __shfl_xor(pi, 16, 64);
        pr += __shfl_xor(pr, 32, 64);
        pi += __shfl_xor(pi, 32, 64);
        if (kg == 0) {
            d.psi[e] = make_double2(pr, pi);
            acc += pr * pr + pi * pi;
        }
    }
    lds[wave][lane] = acc;
    __syncthreads();
    if (wave == 0 && kg == 0)
        d.part[(size_t)rtile * Bp + b] = (lds[0][lane] + lds[1][lane]) + (lds[2][lane] + lds[3][lane]);
}

template <bool TK>
static void launch_combine_bt(const Dev& d, hipStream_t s)
{
    dim3 grid(bpad8(d.ntile_e * (d.sn / 16))), block(kThreads);
    if (d.mixed) {
        if (d.m <= 8) hipLaunchKernelGGL((k_combine_b<float, 2, TK>), grid, block, 0, s, d);
        else if (d.m <= 16) hipLaunchKernelGGL((k_combine_b<float, 4, TK>), grid, block, 0, s, d);
        else if (d.m <= 24) hipLaunchKernelGGL((k_combine_b<float, 6, TK>), grid, block, 0, s, d);
        else hipLaunchKernelGGL((k_combine_b<float, 8, TK>), grid, block, 0, s, d);
    } else {
        if (d.m <= 8) hipLaunchKernelGGL((k_combine_b<double, 2, TK>), grid, block, 0, s, d);
        else if (d.m <= 16) hipLaunchKernelGGL((k_combine_b<double, 4, TK>), grid, block, 0, s, d);
        else if (d.m <= 24) hipLaunchKernelGGL((k_combine_b<double, 6, TK>), grid, block, 0, s, d);
        else hipLaunchKernelGGL((k_combine_b<double, 8, TK>), grid, block, 0, s, d);
    }
}

hipError_t launch_combine_b(const Dev& d, hipStream_t s)
{
    if (d.tk) launch_combine_bt<true>(d, s);
    else launch_combine_bt<false>(d, s);
    return hipGetLastError();
}

}  // namespace dq
