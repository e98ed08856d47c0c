// dq_window.hip -- batched mixed-precision SpMM with the Krylov vector rows
// staged in LDS (lanczosClass.cu:618-661 per trajectory).
//
// The Hamiltonian of an aggregate is banded in its site order (a dipole tube
// in ring-major order couples a site only to rings r-3..r+3), so the x rows a
// tile of T consecutive rows references form one short contiguous window
// [lo, lo + rows).  A workgroup owns a tile and a slice of trajectories, copies
// the window into LDS with coalesced 16-byte loads, and then reads every x row
// of its products from LDS instead of gathering it from L2.
#include "dq_device.h"
#include "dq_internal.h"

namespace dq {

static inline int wpad8(int x) { return (x + 7) & ~7; }

// LDS bytes a launch needs: the largest window, and the partials of 16 waves
size_t spmm_win_lds_bytes(int ts, int max_rows)
{
    const size_t w = (size_t)max_rows * 2 * ts * sizeof(float);
    const size_t r = (size_t)16 * 2 * 64 * sizeof(double);
    return w > r ? w : r;
}

// ---------------------------------------------------------------------------
// The 16-row-block SpMM of dq_batched.hip (k_spmm_mfma) with its x rows read
// from the LDS window of the tile instead of gathered from L2.  A block of 16
// rows over its union of columns (the rb = 16 lists: padded to a multiple of
// 16 columns, 16 couplings per union slot) is a dense (16 x U) x (U x 16 NG)
// FP32 product; a K-step of 4 union columns is NG v_mfma_f32_16x16x4_f32
// sharing one A operand, with the low parts of the couplings and the FP64
// sums of every two 16-slot groups as in k_spmm_mfma.  Lane
// c + 16 k loads A[c][k] (one dword) and the
// column of step k (one dword), both buffer loads issued kDepthM steps ahead,
// then the NG floats of real columns NG c .. NG c + NG - 1 of that window row
// with one LDS read.  The MFMA sequence per output element is that of
// k_spmm_mfma<NG>, so both kernels give bitwise the same product.
// ---------------------------------------------------------------------------
constexpr int kDepthM = 4;

template <int NG> struct LdsV;
template <> struct LdsV<2> { using T = float2; };
template <> struct LdsV<4> { using T = float4; };

template <int NG>
__device__ __forceinline__ float lds_at(const typename LdsV<NG>::T& v, int g)
{
    if constexpr (NG == 2) return g == 0 ? v.x : v.y;
    else return g == 0 ? v.x : (g == 1 ? v.y : (g == 2 ? v.z : v.w));
}

// two bfloat16 (the high halves of a and b, truncated) in one dword: a low
__device__ __forceinline__ uint32_t bf16_pair_w(float a, float b)
{
    return __builtin_amdgcn_perm(__float_as_uint(b), __float_as_uint(a), 0x07060302u);
}

template <int NG, bool EXACT>
__global__ __launch_bounds__(1024) void k_spmm_mfma_win(Dev d, int j, int last)
{
    constexpr int TW = 8 * NG;          // trajectories per slice
    constexpr int RW = 2 * TW;          // floats per window row
    constexpr int C4 = RW / 4;          // float4 per window row
    constexpr int TL = NG / 2;          // trajectories per lane
    using LV = typename LdsV<NG>::T;
    typedef float mf4 __attribute__((ext_vector_type(4)));
    typedef short b4 __attribute__((ext_vector_type(4)));
    extern __shared__ float win[];
    const int nsl = d.sn / TW;
    const int logical = xcd_tile(blockIdx.x, gridDim.x);
    if (logical >= d.ntile_p * nsl) return;
    const int tile = logical / nsl;
    const int bs = d.s0 + (logical % nsl) * TW;          // first trajectory slot of the slice
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, nw = blockDim.x >> 6;
    const int Bp = d.Bp;
    const int r0 = tile * d.win_T, r1 = min(d.n, r0 + d.win_T);
    const int slice_nact = d.nact[bs];
    const bool active = r0 < slice_nact;
    const int lo = d.win_lo[2 * tile], nrows = d.win_lo[2 * tile + 1];
    const float2* __restrict__ x = reinterpret_cast<const float2*>(d.V) + (size_t)j * d.vs;
    float2* __restrict__ y = reinterpret_cast<float2*>(d.V) + (size_t)(j + 1) * d.vs;
    if (active) {
        const int total = nrows * C4;
        for (int q0 = tid; q0 < total; q0 += 4 * (int)blockDim.x) {
            float4 v[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int q = q0 + u * (int)blockDim.x;
                const int r = q / C4, c = q - r * C4;
                v[u] = make_float4(0.f, 0.f, 0.f, 0.f);
                if (q < total && lo + r < slice_nact)
                    v[u] = *reinterpret_cast<const float4*>(x + (size_t)(lo + r) * Bp + bs + 2 * c);
            }
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int q = q0 + u * (int)blockDim.x;
                const int r = q / C4, c = q - r * C4;
                if (q < total) *reinterpret_cast<float4*>(win + r * RW + 4 * c) = v[u];
            }
        }
    }
    __syncthreads();
    const int kq = lane >> 4, nc = lane & 15;
    const uint32_t lane_lds = 4u * NG * nc;                 // bytes of this lane's floats in a window row
    const __amdgpu_buffer_rsrc_t ar = __builtin_amdgcn_make_buffer_rsrc((void*)d.rb_val32, (short)0, 0x7fffffff,
                                                                        0x00020000);
    const __amdgpu_buffer_rsrc_t cr = __builtin_amdgcn_make_buffer_rsrc((void*)d.rb_col, (short)0, 0x7fffffff,
                                                                        0x00020000);
    const __amdgpu_buffer_rsrc_t lr = __builtin_amdgcn_make_buffer_rsrc((void*)d.rb_lo16, (short)0, 0x7fffffff,
                                                                        0x00020000);
    const int nblk = (d.n + 15) / 16, bpt = d.win_T / 16;
    const double gc = d.g;
    const float* __restrict__ diag = reinterpret_cast<const float*>(d.diag);
    const int bl = bs + TL * nc;                            // first of this lane's trajectories
    double acc[TL][2];
#pragma unroll
    for (int t = 0; t < TL; ++t) { acc[t][0] = 0.0; acc[t][1] = 0.0; }
    if (active) {
        for (int rb = tile * bpt + wave; rb < min((tile + 1) * bpt, nblk); rb += nw) {
            const int u0 = uniform(d.rb_ptr[rb]), u1 = uniform(d.rb_ptr[rb + 1]);
            // as k_spmm_mfma: FP32 sums of two 16-slot groups, then FP64
            constexpr int kSumGroups = 2;
            int ng = 0;
            mf4 c[NG];
            double s64[NG][4];
#pragma unroll
            for (int g = 0; g < NG; ++g) {
                c[g] = mf4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
                for (int r = 0; r < 4; ++r) s64[g][r] = 0.0;
            }
            if (u1 > u0) {
                const int nk = (u1 - u0) >> 2;                 // K-steps (the union is padded to 16)
                float av[kDepthM];
                int cv[kDepthM];
                auto lo_at = [&](int p) {                      // p: the first slot of a 16-slot group
                    return __builtin_bit_cast(b4, __builtin_amdgcn_raw_buffer_load_b64(lr, 8u * lane, 32 * p, 0));
                };
                b4 alo = lo_at(u0);
#pragma unroll
                for (int q = 0; q < kDepthM; ++q) {
                    const int p = u0 + 4 * min(q, nk - 1);
                    av[q] = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(ar, (uint32_t)(64 * p) + 4u * lane, 0, 0));
                    cv[q] = (int)__builtin_amdgcn_raw_buffer_load_b32(cr, (uint32_t)(4 * (p + kq)), 0, 0);
                    __builtin_amdgcn_sched_barrier(0);
                }
                // kDepthM = 4 K-steps per iteration: one group of 16 union slots
                for (int k0 = 0; k0 < nk; k0 += kDepthM) {
                    const b4 alo_n = lo_at(u0 + 4 * min(k0 + kDepthM, nk - kDepthM));
                    float xq[NG][kDepthM];
#pragma unroll
                    for (int q = 0; q < kDepthM; ++q) {
                        {
                            const LV xv = *reinterpret_cast<const LV*>(
                                reinterpret_cast<const char*>(win) + (uint32_t)(cv[q] - lo) * (uint32_t)(RW * 4) + lane_lds);
#pragma unroll
                            for (int g = 0; g < NG; ++g) {
                                c[g] = __builtin_amdgcn_mfma_f32_16x16x4f32(av[q], lds_at<NG>(xv, g), c[g], 0, 0, 0);
                                xq[g][q] = lds_at<NG>(xv, g);
                            }
                        }
                        if constexpr (EXACT) {
#pragma unroll
                            for (int g = 0; g < NG; ++g) {
#pragma unroll
                                for (int r = 0; r < 4; ++r) s64[g][r] += (double)c[g][r];
                                c[g] = mf4{0.f, 0.f, 0.f, 0.f};
                            }
                        }
                        const int p = u0 + 4 * min(k0 + q + kDepthM, nk - 1);
                        av[q] = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(ar, (uint32_t)(64 * p) + 4u * lane, 0, 0));
                        cv[q] = (int)__builtin_amdgcn_raw_buffer_load_b32(cr, (uint32_t)(4 * (p + kq)), 0, 0);
                        __builtin_amdgcn_sched_barrier(0);
                    }
#pragma unroll
                    for (int g = 0; g < NG; ++g) {
                        const uint32_t lo2 = bf16_pair_w(xq[g][0], xq[g][1]), hi2 = bf16_pair_w(xq[g][2], xq[g][3]);
                        c[g] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(alo, __builtin_bit_cast(b4, (uint2){lo2, hi2}),
                                                                         c[g], 0, 0, 0);
                    }
                    if (EXACT || ++ng == kSumGroups || k0 + kDepthM >= nk) {
                        ng = 0;
#pragma unroll
                        for (int g = 0; g < NG; ++g) {
#pragma unroll
                            for (int r = 0; r < 4; ++r) s64[g][r] += (double)c[g][r];
                            c[g] = mf4{0.f, 0.f, 0.f, 0.f};
                        }
                    }
                    alo = alo_n;
                }
            }
            // rows 4 kq .. 4 kq + 3 of the block, real columns NG nc .. NG nc + NG - 1
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int r = rb * 16 + 4 * kq + i;
                if (r < r1) {
                    const size_t e = (size_t)r * Bp + bl;
                    const float* xw = win + (r - lo) * RW + NG * nc;
#pragma unroll
                    for (int t = 0; t < TL; ++t) {
                        const float xr = xw[2 * t], xm = xw[2 * t + 1];
                        const float dv = diag[e + t];
                        float wr = (float)__builtin_fma(gc, s64[2 * t][i], (double)dv * (double)xr);
                        float wi = (float)__builtin_fma(gc, s64[2 * t + 1][i], (double)dv * (double)xm);
                        if (r >= d.nact[bl + t]) { wr = 0.f; wi = 0.f; }
                        const double dxr = (double)xr, dxm = (double)xm, dr = (double)wr, dm = (double)wi;
                        if (!last) {
                            y[e + t] = make_float2(wr, wi);
                            acc[t][0] = __builtin_fma(dxm, dm, __builtin_fma(dxr, dr, acc[t][0]));
                            acc[t][1] = __builtin_fma(-dxm, dr, __builtin_fma(dxr, dm, acc[t][1]));
                        } else {
                            acc[t][0] = __builtin_fma(dm, dm, __builtin_fma(dr, dr, acc[t][0]));
                        }
                    }
                }
            }
        }
    }
    // per trajectory: the four lane groups kq, then the waves in order
#pragma unroll
    for (int t = 0; t < TL; ++t)
#pragma unroll
        for (int q = 0; q < 2; ++q) {
            acc[t][q] += __shfl_xor(acc[t][q], 16, 64);
            acc[t][q] += __shfl_xor(acc[t][q], 32, 64);
        }
    __syncthreads();                                         // the window is reused for the partials
    double* red = reinterpret_cast<double*>(win);            // [wave][q][t][nc]
    if (kq == 0) {
#pragma unroll
        for (int t = 0; t < TL; ++t) {
            red[((wave * 2 + 0) * TL + t) * 16 + nc] = acc[t][0];
            red[((wave * 2 + 1) * TL + t) * 16 + nc] = acc[t][1];
        }
    }
    __syncthreads();
    if (wave == 0 && kq == 0) {
        for (int q = 0; q < 2; ++q) {
            if (q == 1 && last) break;
#pragma unroll
            for (int t = 0; t < TL; ++t) {
                double v = 0.0;
                for (int w = 0; w < nw; ++w) v += red[((w * 2 + q) * TL + t) * 16 + nc];
                d.part_s[((size_t)q * d.ntile_p + tile) * Bp + bl + t] = v;
            }
        }
    }
}

hipError_t launch_spmm_win(const Dev& d, int j, int last, hipStream_t s)
{
    const int ts = d.win_ts;
    const dim3 grid(wpad8(d.ntile_p * (d.sn / ts))), block(64 * d.win_w);
    const size_t lds = (size_t)d.win_lds;
    if (d.spmm_exact) {
        if (ts == 32) hipLaunchKernelGGL((k_spmm_mfma_win<4, true>), grid, block, lds, s, d, j, last);
        else hipLaunchKernelGGL((k_spmm_mfma_win<2, true>), grid, block, lds, s, d, j, last);
    } else {
        if (ts == 32) hipLaunchKernelGGL((k_spmm_mfma_win<4, false>), grid, block, lds, s, d, j, last);
        else hipLaunchKernelGGL((k_spmm_mfma_win<2, false>), grid, block, lds, s, d, j, last);
    }
    return hipGetLastError();
}

// lift the dynamic LDS limit of the instantiations to `bytes` (once per process)
hipError_t spmm_win_set_lds_limit(size_t bytes)
{
    const void* fns[4] = {reinterpret_cast<const void*>(&k_spmm_mfma_win<2, false>),
                          reinterpret_cast<const void*>(&k_spmm_mfma_win<4, false>),
                          reinterpret_cast<const void*>(&k_spmm_mfma_win<2, true>),
                          reinterpret_cast<const void*>(&k_spmm_mfma_win<4, true>)};
    hipError_t e = hipSuccess;
    for (int k = 0; k < 4 && e == hipSuccess; ++k)
        e = hipFuncSetAttribute(fns[k], hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
    return e;
}

}  // namespace dq
