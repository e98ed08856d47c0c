// dq_reduce.h -- deterministic tile reductions of the batched Lanczos
// iteration, shared by the SpMM kernels that carry them (dq_batched.hip).
#pragma once
#include "dq_device.h"
#include "dq_internal.h"

namespace dq {

// Sum over tiles t < ntile of p[t Bp] (one quantity's tile column, lane =
// trajectory) in one fixed order: 16 virtual waves, virtual wave w adding
// tile w + 16 u + 128 r into accumulator u (eight of them, summed as a tree),
// then the 16 virtual waves' sums in order.  The block's first NW waves (NW a
// divisor of 16) take virtual waves w, w + NW, ..., so every NW gives the same
// bits.  Every thread of the block calls it (one barrier inside); the result
// is valid in wave 0.  lds: 16 x 64 doubles of the caller's shared memory, free
// again once wave 0 has returned.
template <int NW>
__device__ __forceinline__ double tile_sum16(const double* p, int ntile, int Bp, double (*lds)[64])
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (wave < NW) {
        for (int w = wave; w < 16; w += NW) {
            double a[8] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
            int t = w;
            for (; t + 112 < ntile; t += 128) {
#pragma unroll
                for (int u = 0; u < 8; ++u) a[u] += p[(size_t)(t + 16 * u) * Bp];
            }
#pragma unroll
            for (int u = 0; u < 8; ++u)
                if (t + 16 * u < ntile) a[u] += p[(size_t)(t + 16 * u) * Bp];
            lds[w][lane] = ((a[0] + a[1]) + (a[2] + a[3])) + ((a[4] + a[5]) + (a[6] + a[7]));
        }
    }
    __syncthreads();
    double s = 0.0;
    if (wave == 0)
        for (int w = 0; w < 16; ++w) s += lds[w][lane];
    return s;
}

// tile_sum16 of two tile columns p and q at once (16 loads in flight per
// lane), each summed in exactly tile_sum16's order, by NW waves (NW a divisor
// of 16, every wave of the block); lds: 32 x 64 doubles; one barrier inside;
// the sums are valid in wave 0
template <int NW>
__device__ __forceinline__ void tile_sum16_pair(const double* p, const double* q, int ntile, int Bp,
                                                double (*lds)[64], double& sp, double& sq)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (int w = wave; w < 16; w += NW) {
        double a[8] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
        double c[8] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
        int t = w;
        for (; t + 112 < ntile; t += 128) {
            double x[8], y[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                x[u] = p[(size_t)(t + 16 * u) * Bp];
                y[u] = q[(size_t)(t + 16 * u) * Bp];
            }
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                a[u] += x[u];
                c[u] += y[u];
            }
        }
#pragma unroll
        for (int u = 0; u < 8; ++u)
            if (t + 16 * u < ntile) {
                a[u] += p[(size_t)(t + 16 * u) * Bp];
                c[u] += q[(size_t)(t + 16 * u) * Bp];
            }
        lds[w][lane] = ((a[0] + a[1]) + (a[2] + a[3])) + ((a[4] + a[5]) + (a[6] + a[7]));
        lds[16 + w][lane] = ((c[0] + c[1]) + (c[2] + c[3])) + ((c[4] + c[5]) + (c[6] + c[7]));
    }
    __syncthreads();
    sp = 0.0;
    sq = 0.0;
    if (wave == 0)
        for (int v = 0; v < 16; ++v) {
            sp += lds[v][lane];
            sq += lds[16 + v][lane];
        }
}

// Task `task` of the Gram-row reduction of Lanczos iteration j (j >= 1), run
// by the leading blocks of the SpMM launch (gram_task_blocks): quantity q of
// update j-1 (the Gram row <v_l|v_j> as re/im pairs, then |v_j|^2) for one
// group of 64 slots, red[q][b] = sum over the row tiles of part[q][tile][b].
// Tiles past the group's largest n_active (its first slot, slots are sorted)
// hold exact zeros and are not read.  A whole block returns together.
template <int NW>
__device__ __forceinline__ void gram_task(const Dev& d, int j, int task, double (*lds)[64])
{
    const int nq = 2 * j + 1;
    const int g = task / nq, q = task - g * nq;
    if (g >= d.sn / 64) return;                              // padding to a multiple of 8 blocks
    const int b0 = d.s0 + g * 64, b = b0 + (threadIdx.x & 63);
    const int maxn = min(d.n, uniform(d.nact[b0]));
    const int ntile = min(d.ntile_e, (maxn + d.rt_e - 1) / d.rt_e);
    const double v = tile_sum16<NW>(d.part + (size_t)q * d.ntile_e * d.Bp + b, ntile, d.Bp, lds);
    if ((threadIdx.x >> 6) == 0) d.red[(size_t)q * d.Bp + b] = v;
}

}  // namespace dq
