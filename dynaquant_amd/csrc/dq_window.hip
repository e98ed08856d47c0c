// dq_window.hip -- batched mixed-precision SpMM with the Krylov vector rows
// staged in LDS (lanczosClass.cu:618-661 per trajectory).
//
// The Hamiltonian of an aggregate is banded in its site order (a dipole tube
// in ring-major order couples a site only to rings r-3..r+3), so the x rows a
// tile of T consecutive rows references form one short contiguous window.
//
// k_spmm_roll: a persistent workgroup owns a slice of 32 trajectories and a
// strip of consecutive tiles.  It keeps the x rows of the current tile's
// window in an LDS ring (row r in slot r mod RING, 256 bytes per row), loads
// the next tile's new rows into registers while the current tile multiplies
// and stores them after it: every x row is read from memory about once per
// strip, and the loads never stall the matrix cores.  The product runs as
// 4 x 4 x 1 outer products (v_mfma_f32_4x4x1_16b_f32, sixteen 4 x 4 blocks per
// instruction): a row block is 4 consecutive rows over the sorted union of
// their columns (1.36x the multiply-adds of plain CSR on the 12,000-site tube,
// against 2.7x for 16-row blocks); one union column is one K = 1 step with
// A = the 4 couplings of the block's rows and B = x of the slice's 64 real
// columns (re/im of 32 trajectories).  Every output element is the sequence
// c = fma(J_ik, x_k, c) over the union columns in ascending order, the same
// sequence the 4-row VALU SpMM (dq_batched.hip, k_spmm_b with kRB = 4)
// evaluates, so both give bitwise the same product.
//
// Step lists (built on the host, dq_api.cpp): the union columns of a row
// block are walked in batches of four steps (padded with zero couplings at
// slot 0).  A batch is one record of 18 dwords: the couplings [step][row]
// (16 floats), then the ring slots of the four columns as uint16, two per
// dword.  Each lane loads one dword: lanes 4 k + i hold the coupling of row i
// at step k, which the matrix core broadcasts to all 16 blocks (CBSZ = 4,
// ABID = k); lanes 16, 17 hold the slots, read back as wave-uniform values.
#include "dq_device.h"
#include "dq_internal.h"

namespace dq {

static inline int wpad8(int x) { return (x + 7) & ~7; }

constexpr int kDepth = 4;              // batches of four steps in flight per wave
constexpr int kPre = kRollPre;         // float4 per thread held for the next tile's rows

// ring slot of step k of a batch (record `r` as loaded by the wave)
__device__ __forceinline__ int roll_slot(uint32_t r, int k)
{
    const uint32_t w = (uint32_t)__builtin_amdgcn_readlane((int)r, 16 + (k >> 1));
    return (int)((k & 1) ? (w >> 16) : (w & 0xffffu));
}

__global__ __launch_bounds__(1024) void k_spmm_roll(Dev d, int j, int last)
{
    constexpr int TS = 32, RW = 64, C4 = 16;   // trajectories, floats, float4 per ring row
    typedef float mf4 __attribute__((ext_vector_type(4)));
    extern __shared__ float ring[];
    const int nsl = d.sn / TS;
    const int logical = xcd_tile(blockIdx.x, gridDim.x);
    if (logical >= d.win_strips * nsl) return;
    // strip-major: the slices of a strip (sharing its step lists) run on one XCD
    const int strip = logical / nsl;
    const int b0 = d.s0 + (logical % nsl) * TS;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, nw = blockDim.x >> 6;
    const int nthr = blockDim.x;
    const int Bp = d.Bp, T = d.win_T, RING = d.win_ring;
    const int ntile = (d.n + T - 1) / T, nblk = (d.n + 3) >> 2;
    // slots are ordered by n_active (largest first): the slice's first slot
    // bounds the rows any of its trajectories uses
    const int slice_nact = d.nact[b0];
    const int t0 = strip * d.win_tps;
    const int tact = min(min(ntile, t0 + d.win_tps), (slice_nact + T - 1) / T);
    const float2* __restrict__ x = reinterpret_cast<const float2*>(d.V) + (size_t)j * d.vs;
    float* __restrict__ y = reinterpret_cast<float*>(reinterpret_cast<float2*>(d.V) + (size_t)(j + 1) * d.vs);
    const float* __restrict__ diag = reinterpret_cast<const float*>(d.diag);
    const __amdgpu_buffer_rsrc_t rr = __builtin_amdgcn_make_buffer_rsrc((void*)d.win_rec, (short)0, 0x7fffffff, 0x00020000);
    constexpr uint32_t kRec = 18 * 4;                        // bytes per batch record
    const uint32_t ldw = 4u * (uint32_t)(lane < 16 ? lane : 16 + (lane & 1));
    const uint32_t lane_b = 4u * (uint32_t)lane;             // this lane's real column in a ring row
    const float gc = (float)d.g;
    // x rows at or beyond the slice's n_active are zero in every slot
    auto xrow4 = [&](int row, int c) {
        return row < slice_nact ? *reinterpret_cast<const float4*>(x + (size_t)row * Bp + b0 + 2 * c)
                                : make_float4(0.f, 0.f, 0.f, 0.f);
    };
    double acc0 = 0.0, acc1 = 0.0;
    if (t0 < tact) {
        // the first tile's whole window [lo', hi'], four loads in flight per thread
        {
            const int f0 = d.win_lo[2 * t0], cnt = (d.win_lo[2 * t0 + 1] - f0 + 1) * C4;
            for (int q0 = tid; q0 < cnt; q0 += 4 * nthr) {
                float4 v[4];
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const int q = q0 + u * nthr;
                    v[u] = q < cnt ? xrow4(f0 + q / C4, q % C4) : make_float4(0.f, 0.f, 0.f, 0.f);
                }
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const int q = q0 + u * nthr;
                    if (q < cnt) *reinterpret_cast<float4*>(ring + ((f0 + q / C4) % RING) * RW + 4 * (q % C4)) = v[u];
                }
            }
        }
        __syncthreads();
        for (int t = t0; t < tact; ++t) {
            // the next tile's new rows (hi'_t, hi'_{t+1}] are loaded now and
            // stored after this tile's products; a jump of the window beyond
            // kPre float4 per thread (a wall boundary) loads the rest then
            const int n0 = d.win_lo[2 * t + 1] + 1;
            const int ncnt = t + 1 < tact ? (d.win_lo[2 * (t + 1) + 1] - n0 + 1) * C4 : 0;
            float4 pre[kPre];
#pragma unroll
            for (int u = 0; u < kPre; ++u) {
                const int q = tid + u * nthr;
                pre[u] = q < ncnt ? xrow4(n0 + q / C4, q % C4) : make_float4(0.f, 0.f, 0.f, 0.f);
            }
            const int rb1 = min((t + 1) * (T >> 2), nblk);
            for (int rb = t * (T >> 2) + wave; rb < rb1; rb += nw) {
                const int s0 = uniform(d.win_gptr[rb]), s1 = uniform(d.win_gptr[rb + 1]);
                mf4 c = {0.f, 0.f, 0.f, 0.f};
                if (s1 > s0) {
                    // records kDepth batches ahead; the LDS reads of batch u + 1
                    // before the products of batch u
                    const int bb = s0 >> 2, nb = (s1 - s0) >> 2;
                    uint32_t rec[kDepth];
#pragma unroll
                    for (int u = 0; u < kDepth; ++u) {
                        rec[u] = __builtin_amdgcn_raw_buffer_load_b32(rr, (uint32_t)(bb + min(u, nb - 1)) * kRec + ldw, 0, 0);
                        __builtin_amdgcn_sched_barrier(0);
                    }
                    float bv[2][4];
#pragma unroll
                    for (int k = 0; k < 4; ++k)
                        bv[0][k] = *reinterpret_cast<const float*>(reinterpret_cast<const char*>(ring) +
                                                                   ((uint32_t)roll_slot(rec[0], k) << 8) + lane_b);
                    for (int u = 0; u < nb; u += kDepth) {
#pragma unroll
                        for (int w = 0; w < kDepth; ++w) {
                            const int cur = u + w;
                            if (cur < nb) {
                                if (cur + 1 < nb) {
#pragma unroll
                                    for (int k = 0; k < 4; ++k)
                                        bv[(w + 1) & 1][k] = *reinterpret_cast<const float*>(
                                            reinterpret_cast<const char*>(ring) +
                                            ((uint32_t)roll_slot(rec[(w + 1) % kDepth], k) << 8) + lane_b);
                                }
                                const float a = __uint_as_float(rec[w]);
                                c = __builtin_amdgcn_mfma_f32_4x4x1f32(a, bv[w & 1][0], c, 4, 0, 0);
                                c = __builtin_amdgcn_mfma_f32_4x4x1f32(a, bv[w & 1][1], c, 4, 1, 0);
                                c = __builtin_amdgcn_mfma_f32_4x4x1f32(a, bv[w & 1][2], c, 4, 2, 0);
                                c = __builtin_amdgcn_mfma_f32_4x4x1f32(a, bv[w & 1][3], c, 4, 3, 0);
                            }
                            rec[w] = __builtin_amdgcn_raw_buffer_load_b32(
                                rr, (uint32_t)(bb + min(cur + kDepth, nb - 1)) * kRec + ldw, 0, 0);
                            __builtin_amdgcn_sched_barrier(0);
                        }
                    }
                }
                // rows 4 rb .. 4 rb + 3 of this lane's real column
                const int b = b0 + (lane >> 1);
                const int nact_b = d.nact[b];
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    const int row = 4 * rb + i;
                    const bool valid = row < d.n;
                    float xv = 0.f, wv = 0.f;
                    if (valid) {
                        xv = ring[(row % RING) * RW + lane];
                        const float dv = diag[(size_t)row * Bp + b];
                        wv = __builtin_fmaf(gc, c[i], dv * xv);
                        if (row >= nact_b) wv = 0.f;
                        if (!last) y[((size_t)row * Bp + b0) * 2 + lane] = wv;
                    }
                    // the partner lane holds the other part of the same trajectory
                    const float xp = __shfl_xor(xv, 1, 64), wp = __shfl_xor(wv, 1, 64);
                    if (valid && (lane & 1) == 0) {
                        const double xr = (double)xv, xm = (double)xp, dr = (double)wv, dm = (double)wp;
                        if (!last) {
                            acc0 = __builtin_fma(xm, dm, __builtin_fma(xr, dr, acc0));
                            acc1 = __builtin_fma(-xm, dr, __builtin_fma(xr, dm, acc1));
                        } else {
                            acc0 = __builtin_fma(dm, dm, __builtin_fma(dr, dr, acc0));
                        }
                    }
                }
            }
            __syncthreads();                                 // every wave is done with tile t
#pragma unroll
            for (int u = 0; u < kPre; ++u) {
                const int q = tid + u * nthr;
                if (q < ncnt) *reinterpret_cast<float4*>(ring + ((n0 + q / C4) % RING) * RW + 4 * (q % C4)) = pre[u];
            }
            for (int q = tid + kPre * nthr; q < ncnt; q += nthr)
                *reinterpret_cast<float4*>(ring + ((n0 + q / C4) % RING) * RW + 4 * (q % C4)) = xrow4(n0 + q / C4, q % C4);
            __syncthreads();
        }
    }
    // per trajectory over the waves, in order (the ring is free now)
    __syncthreads();
    double* red = reinterpret_cast<double*>(ring);
    red[(2 * wave) * 64 + lane] = acc0;
    red[(2 * wave + 1) * 64 + lane] = acc1;
    __syncthreads();
    if (wave == 0 && (lane & 1) == 0) {
        const int t = lane >> 1;
        for (int q = 0; q < 2; ++q) {
            if (q == 1 && last) break;
            double v = 0.0;
            for (int w = 0; w < nw; ++w) v += red[(2 * w + q) * 64 + lane];
            d.part_s[((size_t)q * d.ntile_p + strip) * Bp + b0 + t] = v;
        }
    }
}

// LDS bytes a launch needs: the largest window (or ring), and the partials of 16 waves
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
// sharing one A operand.  Lane c + 16 k loads A[c][k] (one dword) and the
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

template <int NG>
__global__ __launch_bounds__(1024) void k_spmm_mfma_win(Dev d, int j, int last)
{
    constexpr int TW = 8 * NG;          // trajectories per slice
    constexpr int RW = 2 * TW;          // floats per window row
    constexpr int C4 = RW / 4;          // float4 per window row
    constexpr int TL = NG / 2;          // trajectories per lane
    using LV = typename LdsV<NG>::T;
    typedef float mf4 __attribute__((ext_vector_type(4)));
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
    const int nblk = (d.n + 15) / 16, bpt = d.win_T / 16;
    const float gc = (float)d.g;
    const float* __restrict__ diag = reinterpret_cast<const float*>(d.diag);
    const int bl = bs + TL * nc;                            // first of this lane's trajectories
    double acc[TL][2];
#pragma unroll
    for (int t = 0; t < TL; ++t) { acc[t][0] = 0.0; acc[t][1] = 0.0; }
    if (active) {
        for (int rb = tile * bpt + wave; rb < min((tile + 1) * bpt, nblk); rb += nw) {
            const int u0 = uniform(d.rb_ptr[rb]), u1 = uniform(d.rb_ptr[rb + 1]);
            mf4 c[NG];
#pragma unroll
            for (int g = 0; g < NG; ++g) c[g] = mf4{0.f, 0.f, 0.f, 0.f};
            if (u1 > u0) {
                const int nk = (u1 - u0) >> 2;                 // K-steps (the union is padded to 16)
                float av[kDepthM];
                int cv[kDepthM];
#pragma unroll
                for (int q = 0; q < kDepthM; ++q) {
                    const int p = u0 + 4 * min(q, nk - 1);
                    av[q] = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(ar, (uint32_t)(64 * p) + 4u * lane, 0, 0));
                    cv[q] = (int)__builtin_amdgcn_raw_buffer_load_b32(cr, (uint32_t)(4 * (p + kq)), 0, 0);
                    __builtin_amdgcn_sched_barrier(0);
                }
                for (int k0 = 0; k0 < nk; k0 += kDepthM) {
#pragma unroll
                    for (int q = 0; q < kDepthM; ++q) {
                        if (k0 + q < nk) {
                            const LV xv = *reinterpret_cast<const LV*>(
                                reinterpret_cast<const char*>(win) + (uint32_t)(cv[q] - lo) * (uint32_t)(RW * 4) + lane_lds);
#pragma unroll
                            for (int g = 0; g < NG; ++g)
                                c[g] = __builtin_amdgcn_mfma_f32_16x16x4f32(av[q], lds_at<NG>(xv, g), c[g], 0, 0, 0);
                        }
                        const int p = u0 + 4 * min(k0 + q + kDepthM, nk - 1);
                        av[q] = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(ar, (uint32_t)(64 * p) + 4u * lane, 0, 0));
                        cv[q] = (int)__builtin_amdgcn_raw_buffer_load_b32(cr, (uint32_t)(4 * (p + kq)), 0, 0);
                        __builtin_amdgcn_sched_barrier(0);
                    }
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
                        float wr = __builtin_fmaf(gc, c[2 * t][i], dv * xr);
                        float wi = __builtin_fmaf(gc, c[2 * t + 1][i], dv * xm);
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
    const size_t lds = (size_t)d.win_lds;
    const dim3 block(64 * d.win_w);
    if (d.rb == 16) {
        // 16-row blocks on 16 x 16 x 4 MFMAs, one tile window per workgroup
        const dim3 grid(wpad8(d.ntile_p * (d.sn / ts)));
        if (ts == 32) hipLaunchKernelGGL((k_spmm_mfma_win<4>), grid, block, lds, s, d, j, last);
        else hipLaunchKernelGGL((k_spmm_mfma_win<2>), grid, block, lds, s, d, j, last);
    } else {
        // 4-row blocks, persistent over strips with the LDS ring
        const dim3 grid(wpad8(d.win_strips * (d.sn / 32)));
        hipLaunchKernelGGL(k_spmm_roll, grid, block, lds, s, d, j, last);
    }
    return hipGetLastError();
}

// lift the dynamic LDS limit of both instantiations to `bytes` (once per process)
hipError_t spmm_win_set_lds_limit(size_t bytes)
{
    const void* fns[3] = {reinterpret_cast<const void*>(&k_spmm_roll),
                          reinterpret_cast<const void*>(&k_spmm_mfma_win<2>),
                          reinterpret_cast<const void*>(&k_spmm_mfma_win<4>)};
    hipError_t e = hipSuccess;
    for (int k = 0; k < 3 && e == hipSuccess; ++k)
        e = hipFuncSetAttribute(fns[k], hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
    return e;
}

}  // namespace dq
