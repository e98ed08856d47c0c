// dq_kernels.hip -- gfx950 kernels of the stochastic exciton-dynamics engine.
//
// One Krylov cycle for all trajectories of a batch (the reference runs one
// trajectory with ~10m+25 cuBLAS/cuSPARSE calls and m+5 host syncs per
// step, propagate.cu:199-260, lanczosClass.cu:405-526):
//
//   prologue        renormalise psi, observables, OU noise -> diagonal,
//                   v_0 = psi                       (systemClass.cu:1769-1971)
//   for j < m:
//     spmm(j)       w = H v_j into slot j+1, partial <v_j|w>
//     scalars_h(j)  beta_{j-1}, scale s_j, breakdown, alpha_j and the
//                   projections h_k = q_k^H H q_j from the Gram row of q_j
//     update(j)     v_{j+1} = s_j w - sum_k h_k s_k v_k, partial Gram row
//                   <v_l|v_{j+1}> and |v_{j+1}|^2
//     reduce        sum the Gram partials over row tiles
//   spmm(last)      |H v_m|^2 for the Expokit estimate (lanczosClass.cu:536-570)
//   pade            exp(-i theta Tbar) per trajectory in LDS (padeClass.cu:204-782)
//   combine         psi = sum_k y_k s_k v_k, partial |psi|^2 (propStep.cu:31-44)
//   finish_norm     |psi|^2 per trajectory
//
// The basis is kept unnormalised; slot k carries a per-trajectory scale s_k
// so normalisation never costs a pass.  Full reorthogonalisation uses the
// identity q_k^H H q_j = (H q_k)^H q_j with H q_k = sum_l h^(k)_l q_l +
// beta_k q_{k+1}: the Gram row of each new vector, accumulated in FP64 in
// the same pass that writes it, gives every projection of the next product
// without re-reading the basis.  All reductions run in a fixed order, so a
// run is bitwise reproducible.
#include "dq_device.h"
#include "dq_internal.h"

namespace dq {

template <typename VT> struct Vec2;
template <> struct Vec2<float> { using T = float2; };
template <> struct Vec2<double> { using T = double2; };

static inline int pad8(int x) { return (x + 7) & ~7; }

// ---------------------------------------------------------------------------
// Block partial sums.  PLAIN: all 256 threads belong to one trajectory and
// the block reduces to one value per quantity.  Batched: lane = trajectory,
// the four waves are summed in LDS.  Quantities go four at a time.
// part layout [q][tile][Bp].
// ---------------------------------------------------------------------------
template <bool PLAIN, int NQ>
__device__ __forceinline__ void store_partials(const double (&acc)[NQ], int nq, double* part,
                                               int ntile, int tile, int Bp, int b)
{
    __shared__ double lds[4][4][64];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
#pragma unroll
    for (int q0 = 0; q0 < NQ; q0 += 4) {
        if (q0 < nq) {
            __syncthreads();
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                if (q0 + u < NQ) {
                    double v = acc[q0 + u < NQ ? q0 + u : 0];
                    if (PLAIN) {
                        v = wave_sum(v);
                        if (lane == 0) lds[u][wave][0] = v;
                    } else {
                        lds[u][wave][lane] = v;
                    }
                }
            }
            __syncthreads();
            if (wave == 0) {
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    int q = q0 + u;
                    if (q < nq && q < NQ) {
                        if (PLAIN) {
                            if (lane == 0)
                                part[(size_t)q * ntile + tile] =
                                    (lds[u][0][0] + lds[u][1][0]) + (lds[u][2][0] + lds[u][3][0]);
                        } else {
                            part[((size_t)q * ntile + tile) * Bp + b] =
                                (lds[u][0][lane] + lds[u][1][lane]) + (lds[u][2][lane] + lds[u][3][lane]);
                        }
                    }
                }
            }
        }
    }
}

// Sum part[q][0..ntile)[b] inside one block; every thread gets the result.
// PLAIN: 256 threads share trajectory 0.  Batched: lane = trajectory, the
// waves split the tiles.
template <bool PLAIN>
__device__ __forceinline__ double block_tile_sum(const double* part, int q, int ntile, int Bp, int b)
{
    __shared__ double lds[4][64];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    double s = 0.0;
    __syncthreads();
    if (PLAIN) {
        for (int t = tid; t < ntile; t += kThreads) s += part[(size_t)q * ntile + t];
        s = wave_sum(s);
        if (lane == 0) lds[wave][0] = s;
        __syncthreads();
        s = (lds[0][0] + lds[1][0]) + (lds[2][0] + lds[3][0]);
    } else {
        // eight independent loads in flight per lane, fixed summation order
        double a[8] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
        const double* p = part + (size_t)q * ntile * Bp + b;
        int t = wave;
        for (; t + 28 < ntile; t += 32) {
#pragma unroll
            for (int u = 0; u < 8; ++u) a[u] += p[(size_t)(t + 4 * u) * Bp];
        }
#pragma unroll
        for (int u = 0; u < 8; ++u)
            if (t + 4 * u < ntile) a[u] += p[(size_t)(t + 4 * u) * Bp];
        s = ((a[0] + a[1]) + (a[2] + a[3])) + ((a[4] + a[5]) + (a[6] + a[7]));
        lds[wave][lane] = s;
        __syncthreads();
        s = (lds[0][lane] + lds[1][lane]) + (lds[2][lane] + lds[3][lane]);
    }
    __syncthreads();
    return s;
}

// elementwise tile decomposition shared by prologue, update and combine
struct Tile {
    int rtile, b, r0, r1, rstep, rfirst;
    bool ok;
};

// tile of a workgroup over rows in tiles of rt (ntile of them) x 64 trajectories
template <bool PLAIN>
__device__ __forceinline__ Tile elem_tile(const Dev& d, int rt, int ntile)
{
    Tile t;
    const int ntt = PLAIN ? 1 : d.sn / 64;
    const int nblk = ntile * ntt;
    const int logical = xcd_tile(blockIdx.x, gridDim.x);
    t.ok = logical < nblk;
    t.rtile = logical / ntt;
    const int ttile = (PLAIN ? 0 : d.s0 / 64) + logical % ntt;
    const int tid = threadIdx.x;
    t.b = PLAIN ? 0 : ttile * 64 + (tid & 63);
    t.r0 = t.rtile * rt;
    t.r1 = min(d.n, t.r0 + rt);
    t.rstep = PLAIN ? kThreads : 4;
    t.rfirst = t.r0 + (PLAIN ? tid : (tid >> 6));
    return t;
}

static inline int elem_grid(const Dev& d, int ntile) { return pad8(ntile * (d.plain ? 1 : d.sn / 64)); }

// ---------------------------------------------------------------------------
// Static disorder and stationary OU start (systemClass.cu:759-804, 1028-1036)
// ---------------------------------------------------------------------------
__global__ void k_init_disorder(Dev d)
{
#pragma clang fp contract(off)
    const size_t total = (size_t)d.n * d.Bp;
    for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < total;
         e += (size_t)gridDim.x * blockDim.x) {
        int i = (int)(e / d.Bp), b = (int)(e % d.Bp);
        double e0 = 0.0, de = 0.0;
        if (b < d.B && i < d.nact[b]) {
            double z[4];
            double base = d.e_site[i];
            if (d.sigma_static > 0.0) {
                normal4(d.seed, d.traj_id[b], (uint64_t)i, 0, kStreamStatic, 0, z);
                e0 = d.sigma_static * z[0] + base;
            } else {
                e0 = base;
            }
            if (d.noise_model == 0) {               // stationary OU start (systemClass.cu:1036)
                normal4(d.seed, d.traj_id[b], (uint64_t)i, 0, kStreamOuInit, 0, z);
                de = d.ou_sigma[i] * z[0];
            }
        }
        d.e0[e] = e0;
        d.de[e] = de;
        if (d.mixed) reinterpret_cast<float*>(d.diag)[e] = (float)(de + e0);
        else reinterpret_cast<double*>(d.diag)[e] = de + e0;
    }
}

hipError_t launch_init_disorder(const Dev& d, hipStream_t s)
{
    size_t total = (size_t)d.n * d.Bp;
    int grid = (int)((total + kThreads - 1) / kThreads);
    if (grid > 65535) grid = 65535;
    hipLaunchKernelGGL(k_init_disorder, dim3(grid), dim3(kThreads), 0, s, d);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Prologue of a Krylov cycle: renormalise (systemClass.cu:1769-1829),
// record observables of psi(step) (popPrintClass.cu:442-472,
// spectraClass.cu:509-571), advance the OU noise of this step and write the
// diagonal (gpuSysHelper.cu:73-128), copy psi into basis slot 0.
// ---------------------------------------------------------------------------
template <bool PLAIN, typename VT>
__global__ __launch_bounds__(kThreads) void k_prologue(Dev d, int renorm, int noise, int record, int copy_v0)
{
#pragma clang fp contract(off)
    using V2 = typename Vec2<VT>::T;
    Tile t = elem_tile<PLAIN>(d, d.rt_o, d.ntile_o);
    if (!t.ok) return;
    const int b = t.b;
    const int64_t step = *d.step;
    int rec_idx = -1;
    if (record && d.every > 0 && step % d.every == 0 && step <= d.max_steps) rec_idx = (int)(step / d.every);
    const bool corr_on = record && d.bra != nullptr && step <= d.max_steps;
    const bool pr_on = record && d.partic != nullptr && step <= d.max_steps;
    const bool grp = rec_idx >= 0 && !d.site_mode;
    double inv = 1.0;
    if (renorm) {
        double n2 = d.nrm2[b];
        inv = n2 > 0.0 ? 1.0 / __builtin_sqrt(n2) : 0.0;
    }
    const int nact = d.nact[b];
    const uint64_t tid_b = d.traj_id[b];
    // this trajectory's bra (a table of bras, e.g. one per dipole direction)
    const double2* __restrict__ brow = corr_on ? d.bra + (d.bra_idx ? (size_t)d.bra_idx[b] * d.n : 0) : nullptr;
    double acc[kMaxGroups + 3];
#pragma unroll
    for (int q = 0; q < kMaxGroups + 3; ++q) acc[q] = 0.0;
    V2* v0 = reinterpret_cast<V2*>(d.V);
    VT* diag = reinterpret_cast<VT*>(d.diag);
    // trajectory slots are sorted by n_active (largest first), so rows at or
    // beyond the slice's first n_active are inactive for all of its lanes
    const int rend = PLAIN ? t.r1 : min(t.r1, (int)uniform(d.nact[b & ~63]));
    for (int i = t.rfirst; i < rend; i += t.rstep) {
        const size_t e = (size_t)i * d.Bp + b;
        double2 p = d.psi[e];
        if (renorm) {
            p.x = p.x * inv;
            p.y = p.y * inv;
            d.psi[e] = p;
        }
        if (rec_idx >= 0) {
            double pp = p.x * p.x + p.y * p.y;
            if (d.site_mode) {
                d.pops[((size_t)rec_idx * d.n + i) * d.Bp + b] = pp;
            } else {
                int g = d.group[i];
#pragma unroll
                for (int q = 0; q < kMaxGroups; ++q)
                    if (g == q) acc[q] += pp;
            }
        }
        if (corr_on) {
            double2 br = brow[i];
            acc[kMaxGroups] += br.x * p.x + br.y * p.y;
            acc[kMaxGroups + 1] += br.x * p.y - br.y * p.x;
        }
        if (pr_on) {
            // sum_i |c_i|^4 (systemClass.cu:2209-2215)
            double pp = p.x * p.x + p.y * p.y;
            acc[kMaxGroups + 2] += pp * pp;
        }
        if (noise && i < nact) {
            double cur = d.de[e];
            if (d.noise_model == 0) {
                // Kubo-Anderson / OU, S Euler-Maruyama substeps (gpuSysHelper.cu:73-128)
                const double invTau = d.inv_tau[i], coeff = d.coeff[i];
                if (coeff != 0.0 || invTau != 0.0) {
                    // substep s uses normal s & 3 of Philox block s >> 2; one
                    // block per four substeps, and only the Box-Muller pair in
                    // use is evaluated (the same values normal4 returns)
                    uint64_t c[4] = {0, 0, 0, 0};
                    double z[2] = {0.0, 0.0};
                    for (int s = 0; s < d.S; ++s) {
                        if ((s & 3) == 0) {
                            c[0] = (uint64_t)i; c[1] = (uint64_t)step; c[2] = kStreamOu; c[3] = (uint64_t)(s >> 2);
                            philox4x64_10(c, d.seed, tid_b);
                        }
                        if ((s & 1) == 0) {
                            if ((s & 2) == 0) box_muller(c[0], c[1], &z[0], &z[1]);
                            else box_muller(c[2], c[3], &z[0], &z[1]);
                        }
                        const double xi = z[s & 1];
                        cur = cur - cur * invTau * d.subdt + coeff * xi * d.sqrt_subdt;
                    }
                    d.de[e] = cur;
                }
            } else if (d.noise_model == 1) {
                // Haken-Strobl-Reineker white noise (gpuSysHelper.cu:32-66)
                double z[4];
                normal4(d.seed, tid_b, (uint64_t)i, (uint64_t)step, kStreamHsr, 0, z);
                cur = d.hsr_sigma[i] * z[0];
                d.de[e] = cur;
            } else if (d.noise_model == 2) {
                // Zhong-Zhao real noise: sum_k c_k cos(omega_k t + phi_k) (gpuSysHelper.cu:134-177)
                const double t = (double)step * d.dt;
                double acc = 0.0;
                for (int q = 0; q < d.zz_nosc; q += 4) {
                    double u[4];
                    zz_phase4(d.seed, tid_b, (uint64_t)i, (uint64_t)(q >> 2), u);
                    for (int r = 0; r < 4 && q + r < d.zz_nosc; ++r)
                        acc = acc + d.zz_coeff[q + r] * zz_cos(u[r], d.zz_w2pi[q + r], t);
                }
                cur = acc;
                d.de[e] = cur;
            } else {
                cur = 0.0;                          // constant Hamiltonian
            }
            diag[e] = (VT)(cur + d.e0[e]);
        }
        if (copy_v0) {
            V2 v;
            v.x = (VT)p.x;
            v.y = (VT)p.y;
            v0[e] = v;
        }
    }
    if (grp || corr_on || pr_on) {
        // quantities: group g at q = g, C(n) at q = kMaxGroups, kMaxGroups + 1,
        // sum |c|^4 at q = kMaxGroups + 2
        store_partials<PLAIN, kMaxGroups + 3>(acc, kMaxGroups + 3, d.part, d.ntile_o, t.rtile, d.Bp, b);
    }
}

hipError_t launch_prologue(const Dev& d, int renorm, int noise, int record, int copy_v0, hipStream_t s)
{
    dim3 grid(elem_grid(d, d.ntile_o)), block(kThreads);
    if (d.plain) {
        if (d.mixed) hipLaunchKernelGGL((k_prologue<true, float>), grid, block, 0, s, d, renorm, noise, record, copy_v0);
        else hipLaunchKernelGGL((k_prologue<true, double>), grid, block, 0, s, d, renorm, noise, record, copy_v0);
    } else {
        if (d.mixed) hipLaunchKernelGGL((k_prologue<false, float>), grid, block, 0, s, d, renorm, noise, record, copy_v0);
        else hipLaunchKernelGGL((k_prologue<false, double>), grid, block, 0, s, d, renorm, noise, record, copy_v0);
    }
    return hipGetLastError();
}

// observables of the prologue: partial sums -> pops[rec][g][b], corr[step][b],
// partic[step][b]
template <bool PLAIN>
__global__ __launch_bounds__(kThreads) void k_reduce_obs(Dev d)
{
    const int64_t step = *d.step;
    int rec_idx = -1;
    if (d.every > 0 && step % d.every == 0 && step <= d.max_steps) rec_idx = (int)(step / d.every);
    const bool corr_on = d.bra != nullptr && step <= d.max_steps;
    const bool pr_on = d.partic != nullptr && step <= d.max_steps;
    const bool grp = rec_idx >= 0 && !d.site_mode;
    if (!grp && !corr_on && !pr_on) return;
    const int b = PLAIN ? 0 : d.s0 + blockIdx.x * 64 + (threadIdx.x & 63);
    if (grp) {
        for (int g = 0; g < d.ngroups; ++g) {
            double s = block_tile_sum<PLAIN>(d.part, g, d.ntile_o, d.Bp, b);
            if (threadIdx.x < (PLAIN ? 1 : 64)) d.pops[((size_t)rec_idx * d.ngroups + g) * d.Bp + b] = s;
        }
    }
    if (corr_on) {
        double re = block_tile_sum<PLAIN>(d.part, kMaxGroups, d.ntile_o, d.Bp, b);
        double im = block_tile_sum<PLAIN>(d.part, kMaxGroups + 1, d.ntile_o, d.Bp, b);
        if (threadIdx.x < (PLAIN ? 1 : 64)) d.corr[(size_t)step * d.Bp + b] = make_double2(re, im);
    }
    if (pr_on) {
        // the reference stores 1 / sum |c|^4 (systemClass.cu:2251-2252)
        double s4 = block_tile_sum<PLAIN>(d.part, kMaxGroups + 2, d.ntile_o, d.Bp, b);
        if (threadIdx.x < (PLAIN ? 1 : 64)) d.partic[(size_t)step * d.Bp + b] = s4 > 0.0 ? 1.0 / s4 : 0.0;
    }
}

hipError_t launch_reduce_obs(const Dev& d, hipStream_t s)
{
    if (d.plain) hipLaunchKernelGGL((k_reduce_obs<true>), dim3(1), dim3(kThreads), 0, s, d);
    else hipLaunchKernelGGL((k_reduce_obs<false>), dim3(d.sn / 64), dim3(kThreads), 0, s, d);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// SpMV of the single-trajectory path; the batched SpMM is in dq_batched.hip.
// ---------------------------------------------------------------------------
// Single-trajectory CSR SpMV (FP64): 16 lanes per row, 64 rows per block.
__global__ __launch_bounds__(kThreads) void k_spmv_plain(Dev d, int j, int last)
{
    const int rtile = xcd_tile(blockIdx.x, gridDim.x);
    if (rtile >= d.ntile_s) return;
    const int tid = threadIdx.x, grp = tid >> 4, l = tid & 15;
    const size_t nB = (size_t)d.n;
    const double2* __restrict__ x = reinterpret_cast<const double2*>(d.V) + (size_t)j * d.vs;
    double2* __restrict__ y = reinterpret_cast<double2*>(d.V) + (size_t)(j + 1) * d.vs;
    const double* __restrict__ diag = reinterpret_cast<const double*>(d.diag);
    const int nact = d.nact[0];
    double acc0 = 0.0, acc1 = 0.0;
    const int r0 = rtile * d.rt_s, r1 = min(d.n, r0 + d.rt_s);
    for (int i = r0 + grp; i < r1; i += 16) {
        const int p0 = d.rowptr[i], p1 = d.rowptr[i + 1];
        double sr = 0.0, si = 0.0;
        for (int p = p0 + l; p < p1; p += 16) {
            const double a = d.val64[p];
            const double2 xv = x[d.col[p]];
            sr += a * xv.x;
            si += a * xv.y;
        }
        sr = group_sum(sr, 16);
        si = group_sum(si, 16);
        if (l == 0) {
            const double2 xi = x[i];
            const double di = diag[i];
            double wr = di * xi.x + d.g * sr;
            double wi = di * xi.y + d.g * si;
            if (i >= nact) { wr = 0.0; wi = 0.0; }
            if (!last) {
                y[i] = make_double2(wr, wi);
                acc0 += xi.x * wr + xi.y * wi;
                acc1 += xi.x * wi - xi.y * wr;
            } else {
                acc0 += wr * wr + wi * wi;
            }
        }
    }
    double out[2] = {acc0, acc1};
    store_partials<true, 2>(out, last ? 1 : 2, d.part, d.ntile_s, rtile, 1, 0);
}

hipError_t launch_spmm(const Dev& d, int j, int last, hipStream_t s)
{
    if (!d.plain) return launch_spmm_b(d, j, last, s);
    hipLaunchKernelGGL(k_spmv_plain, dim3(pad8(d.ntile_s)), dim3(kThreads), 0, s, d, j, last);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Per-trajectory scalars of Lanczos iteration j.
//   j >= 1: beta_{j-1} = |v_j| from the reduced Gram sums of update(j-1),
//           s_j = 1/beta_{j-1} or 0 on breakdown (beta_{j-1} <= tol
//           (|alpha_{j-1}| + beta_{j-2})), Gram row G_l = <q_l|q_j>.
//   h_j = s_j^2 <v_j|H v_j>,  alpha_j = Re h_j  (Hermitian alpha, the
//           reference uses the unconjugated cublasZdotu, lanczosClass.cu:421)
//   h_k = beta_k G_{k+1} + sum_{l<=k} conj(h^(k)_l) G_l   (k < j)
// PLAIN: one block, threads split k.  Batched: lane = trajectory, waves
// split k.
// ---------------------------------------------------------------------------
template <bool PLAIN>
__global__ __launch_bounds__(kThreads) void k_scalars_h(Dev d, int j)
{
    __shared__ double s_sj[64];
    __shared__ double2 s_G[kMaxM + 1][64];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int bl = PLAIN ? 0 : lane;
    const int b = PLAIN ? 0 : blockIdx.x * 64 + lane;
    const int Bp = d.Bp;
    // <v_j|w> summed over the SpMM tiles
    const double cr = block_tile_sum<PLAIN>(d.part, 0, d.ntile_s, Bp, b);
    const double ci = block_tile_sum<PLAIN>(d.part, 1, d.ntile_s, Bp, b);
    const bool writer = PLAIN ? (tid == 0) : (wave == 0);
    if (writer) {
        double sj = 1.0;
        if (j == 0) {
            s_G[0][bl] = make_double2(1.0, 0.0);
            d.sc[b] = 1.0;
            d.broke[b] = 0;
        } else {
            const double nu = d.red[(size_t)(2 * j) * Bp + b];
            double bet = __builtin_sqrt(nu);
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
            d.broke[b] = broke;
            d.beta[(size_t)(j - 1) * Bp + b] = bet;
            d.sc[(size_t)j * Bp + b] = sj;
            for (int l = 0; l < j; ++l) {
                const double sl = d.sc[(size_t)l * Bp + b];
                const double f = sl * sj;
                double2 G = make_double2(f * d.red[(size_t)(2 * l) * Bp + b],
                                         f * d.red[(size_t)(2 * l + 1) * Bp + b]);
                s_G[l][bl] = G;
                d.gram[(size_t)l * Bp + b] = G;
            }
            s_G[j][bl] = make_double2(1.0, 0.0);
        }
        s_sj[bl] = sj;
        const double hr = sj * sj * cr, hi = sj * sj * ci;
        d.alpha[(size_t)j * Bp + b] = hr;
        d.hco[((size_t)j * kMaxM + j) * Bp + b] = make_double2(hr, hi);
    }
    __syncthreads();
    // projections on the older vectors
    const int kstart = PLAIN ? tid : wave;
    const int kstep = PLAIN ? kThreads : 4;
    for (int k = kstart; k < j; k += kstep) {
        const double bk = d.beta[(size_t)k * Bp + b];
        double2 Gk1 = s_G[k + 1][bl];
        double sr = bk * Gk1.x, si = bk * Gk1.y;
        for (int l = 0; l <= k; ++l) {
            const double2 a = d.hco[((size_t)k * kMaxM + l) * Bp + b];
            const double2 G = s_G[l][bl];
            sr += a.x * G.x + a.y * G.y;
            si += a.x * G.y - a.y * G.x;
        }
        d.hco[((size_t)j * kMaxM + k) * Bp + b] = make_double2(sr, si);
    }
}

hipError_t launch_scalars_h(const Dev& d, int j, hipStream_t s)
{
    if (!d.plain) return launch_scalars_hb(d, j, s);
    hipLaunchKernelGGL((k_scalars_h<true>), dim3(1), dim3(kThreads), 0, s, d, j);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Update of iteration j: v_{j+1} = s_j w - sum_{k<=j} (h_k s_k) v_k, with the
// FP64 partial Gram row <v_l|v_{j+1}> (l <= j) and |v_{j+1}|^2 of the
// rounded, stored vector.  MAXV >= j+1 bounds the register arrays.
// ---------------------------------------------------------------------------
template <bool PLAIN, typename VT, int MAXV>
__global__ __launch_bounds__(kThreads) void k_update(Dev d, int j)
{
    using V2 = typename Vec2<VT>::T;
    __shared__ double2 s_c[MAXV][64];
    __shared__ double s_w[64];
    Tile t = elem_tile<PLAIN>(d, d.rt_e, d.ntile_e);
    if (!t.ok) return;
    const int b = t.b, bl = PLAIN ? 0 : (threadIdx.x & 63);
    const int Bp = d.Bp;
    // coefficients -h_k s_k and s_j for this block's trajectories
    if (PLAIN ? threadIdx.x <= (unsigned)j : threadIdx.x < 64) {
        if (PLAIN) {
            const int k = threadIdx.x;
            const double2 h = d.hco[((size_t)j * kMaxM + k)];
            const double sk = d.sc[k];
            s_c[k][0] = make_double2(-h.x * sk, -h.y * sk);
            if (k == 0) s_w[0] = d.sc[j];
        } else {
            for (int k = 0; k <= j; ++k) {
                const double2 h = d.hco[((size_t)j * kMaxM + k) * Bp + b];
                const double sk = d.sc[(size_t)k * Bp + b];
                s_c[k][bl] = make_double2(-h.x * sk, -h.y * sk);
            }
            s_w[bl] = d.sc[(size_t)j * Bp + b];
        }
    }
    __syncthreads();
    const double sw = s_w[bl];
    const size_t nB = (size_t)d.n * Bp;
    const V2* __restrict__ V = reinterpret_cast<const V2*>(d.V);
    V2* __restrict__ out = reinterpret_cast<V2*>(d.V) + (size_t)(j + 1) * d.vs;
    double acc[2 * MAXV + 1];
#pragma unroll
    for (int q = 0; q < 2 * MAXV + 1; ++q) acc[q] = 0.0;
    for (int i = t.rfirst; i < t.r1; i += t.rstep) {
        const size_t e = (size_t)i * Bp + b;
        V2 vk[MAXV];
#pragma unroll
        for (int k = 0; k < MAXV; ++k)
            if (k <= j) vk[k] = V[(size_t)k * d.vs + e];
        const V2 w = out[e];
        double ar = sw * (double)w.x, ai = sw * (double)w.y;
#pragma unroll
        for (int k = 0; k < MAXV; ++k) {
            if (k <= j) {
                const double2 cf = s_c[k][bl];
                const double xr = (double)vk[k].x, xi = (double)vk[k].y;
                ar += cf.x * xr - cf.y * xi;
                ai += cf.x * xi + cf.y * xr;
            }
        }
        V2 nv;
        nv.x = (VT)ar;
        nv.y = (VT)ai;
        out[e] = nv;
        const double nr = (double)nv.x, ni = (double)nv.y;
#pragma unroll
        for (int k = 0; k < MAXV; ++k) {
            if (k <= j) {
                const double xr = (double)vk[k].x, xi = (double)vk[k].y;
                acc[2 * k] += xr * nr + xi * ni;
                acc[2 * k + 1] += xr * ni - xi * nr;
            }
        }
        acc[2 * MAXV] += nr * nr + ni * ni;
    }
    // quantities: G_0..G_j (re, im) then nu at q = 2(j+1)
    double outq[2 * MAXV + 1];
#pragma unroll
    for (int q = 0; q < 2 * MAXV + 1; ++q) outq[q] = acc[q];
    // move nu next to the last Gram entry
#pragma unroll
    for (int q = 0; q < 2 * MAXV + 1; ++q)
        if (q == 2 * (j + 1)) outq[q] = acc[2 * MAXV];
    store_partials<PLAIN, 2 * MAXV + 1>(outq, 2 * (j + 1) + 1, d.part, d.ntile_e, t.rtile, Bp, b);
}

template <bool PLAIN, typename VT>
static void launch_update_t(const Dev& d, int j, hipStream_t s)
{
    dim3 grid(elem_grid(d, d.ntile_e)), block(kThreads);
    if (d.m <= 8) {
        if (!launch_with_events(k_update<PLAIN, VT, 8>, grid, block, s, d, j))
            hipLaunchKernelGGL((k_update<PLAIN, VT, 8>), grid, block, 0, s, d, j);
    } else if (d.m <= 16) {
        if (!launch_with_events(k_update<PLAIN, VT, 16>, grid, block, s, d, j))
            hipLaunchKernelGGL((k_update<PLAIN, VT, 16>), grid, block, 0, s, d, j);
    } else if (d.m <= 24) {
        if (!launch_with_events(k_update<PLAIN, VT, 24>, grid, block, s, d, j))
            hipLaunchKernelGGL((k_update<PLAIN, VT, 24>), grid, block, 0, s, d, j);
    } else {
        if (!launch_with_events(k_update<PLAIN, VT, 32>, grid, block, s, d, j))
            hipLaunchKernelGGL((k_update<PLAIN, VT, 32>), grid, block, 0, s, d, j);
    }
}

hipError_t launch_update(const Dev& d, int j, hipStream_t s)
{
    if (!d.plain) return launch_update_b(d, j, s);
    launch_update_t<true, double>(d, j, s);
    return hipGetLastError();
}

// red[q][b] = sum over tiles of part[q][tile][b]; grid (Bp/64, nq) or (1, nq)
template <bool PLAIN>
__global__ __launch_bounds__(kThreads) void k_reduce(Dev d, int ntile)
{
    const int q = blockIdx.y;
    const int b = PLAIN ? 0 : d.s0 + blockIdx.x * 64 + (threadIdx.x & 63);
    double s = block_tile_sum<PLAIN>(d.part, q, ntile, d.Bp, b);
    if (threadIdx.x < (PLAIN ? 1u : 64u)) d.red[(size_t)q * d.Bp + b] = s;
}

hipError_t launch_reduce(const Dev& d, int nq, int ntile, hipStream_t s)
{
    if (d.plain) hipLaunchKernelGGL((k_reduce<true>), dim3(1, nq), dim3(kThreads), 0, s, d, ntile);
    else hipLaunchKernelGGL((k_reduce<false>), dim3(d.sn / 64, nq), dim3(kThreads), 0, s, d, ntile);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Pade(6,6) with scaling and squaring of A = -i theta Tbar, in LDS, one
// block per trajectory (padeClass.cu:204-487 with the true infinity norm;
// Tbar is the augmented tridiagonal of gpuLanczosHelper.cu:14-60).
// With Tbar real, A = i a (a = -theta 2^-s Tbar real), so the even powers are
// real: P2 = a a, P4 = P2 P2, P6 = P2 P4;  V = c0 - c2 P2 + c4 P4 - c6 P6,
// U = i a W with W = c1 - c3 P2 + c5 P4;  X = (V - U)^-1 U by Gauss-Jordan
// with partial pivoting (the reference calls culaDeviceZgesv,
// padeClass.cu:714-745);  E = I + 2X squared s times.
// Row-major n x n in LDS; regions X0, X1 (complex), R0, R1 (real).
// ---------------------------------------------------------------------------
constexpr int kNMax = kMaxM + 2;
constexpr int kNN = kNMax * kNMax;
// threads of a Pade block: the Gauss-Jordan elimination of a pivot (n x 2n
// complex elements) spread over 16 waves is a short chain per thread
#ifndef PADE_THREADS
#define PADE_THREADS 1024
#endif
constexpr int kPadeThreads = PADE_THREADS;

// augmented tridiagonal of size m + aug (gpuLanczosHelper.cu:14-60): T, a 1
// at (0, m) and ones on the superdiagonal of the trailing aug x aug block;
// column m + j - 1 of its exponential holds the phi_j block
__device__ __forceinline__ double tbar_at(int i, int jj, int m, int aug, const double* al, const double* be)
{
    if (i < m && jj < m) {
        if (i == jj) return al[i];
        if (jj == i + 1) return be[i];
        if (i == jj + 1) return be[jj];
        return 0.0;
    }
    if (i == 0 && jj == m) return 1.0;
    if (i >= m && jj == i + 1 && jj < m + aug) return 1.0;
    return 0.0;
}

// Products of n x n matrices in LDS, a 2 x 2 block of C per thread (one pass
// for n <= 32: half the LDS reads of one element per thread), the k loop
// unrolled four times so that the loads of four k are in flight together.
// Each element is the same k-ordered sum as a one-element-per-thread product.
__device__ void mm_real(int n, const double* A, const double* B, double* C)
{
    const int nb = (n + 1) >> 1;
    for (int blk = threadIdx.x; blk < nb * nb; blk += blockDim.x) {
        const int i0 = 2 * (blk / nb), j0 = 2 * (blk % nb);
        const int i1 = i0 + 1 < n ? i0 + 1 : i0, j1 = j0 + 1 < n ? j0 + 1 : j0;
        double s00 = 0.0, s01 = 0.0, s10 = 0.0, s11 = 0.0;
#pragma unroll 4
        for (int k = 0; k < n; ++k) {
            const double a0 = A[i0 * n + k], a1 = A[i1 * n + k];
            const double b0 = B[k * n + j0], b1 = B[k * n + j1];
            s00 += a0 * b0;
            s01 += a0 * b1;
            s10 += a1 * b0;
            s11 += a1 * b1;
        }
        C[i0 * n + j0] = s00;
        if (j0 + 1 < n) C[i0 * n + j0 + 1] = s01;
        if (i0 + 1 < n) C[(i0 + 1) * n + j0] = s10;
        if (i0 + 1 < n && j0 + 1 < n) C[(i0 + 1) * n + j0 + 1] = s11;
    }
    __syncthreads();
}

__device__ void mm_cplx(int n, const double2* A, const double2* B, double2* C)
{
    const int nb = (n + 1) >> 1;
    for (int blk = threadIdx.x; blk < nb * nb; blk += blockDim.x) {
        const int i0 = 2 * (blk / nb), j0 = 2 * (blk % nb);
        const int i1 = i0 + 1 < n ? i0 + 1 : i0, j1 = j0 + 1 < n ? j0 + 1 : j0;
        double r00 = 0.0, q00 = 0.0, r01 = 0.0, q01 = 0.0, r10 = 0.0, q10 = 0.0, r11 = 0.0, q11 = 0.0;
#pragma unroll 4
        for (int k = 0; k < n; ++k) {
            const double2 a0 = A[i0 * n + k], a1 = A[i1 * n + k];
            const double2 b0 = B[k * n + j0], b1 = B[k * n + j1];
            r00 += a0.x * b0.x - a0.y * b0.y;
            q00 += a0.x * b0.y + a0.y * b0.x;
            r01 += a0.x * b1.x - a0.y * b1.y;
            q01 += a0.x * b1.y + a0.y * b1.x;
            r10 += a1.x * b0.x - a1.y * b0.y;
            q10 += a1.x * b0.y + a1.y * b0.x;
            r11 += a1.x * b1.x - a1.y * b1.y;
            q11 += a1.x * b1.y + a1.y * b1.x;
        }
        C[i0 * n + j0] = make_double2(r00, q00);
        if (j0 + 1 < n) C[i0 * n + j0 + 1] = make_double2(r01, q01);
        if (i0 + 1 < n) C[(i0 + 1) * n + j0] = make_double2(r10, q10);
        if (i0 + 1 < n && j0 + 1 < n) C[(i0 + 1) * n + j0 + 1] = make_double2(r11, q11);
    }
    __syncthreads();
}

// lane l's value of x, read by every lane of the wave (l uniform)
__device__ __forceinline__ double readlane_f64(double x, int l)
{
    const long long b = __double_as_longlong(x);
    const int lo = __builtin_amdgcn_readlane((int)b, l);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// Returns the scaling exponent s (or -1 when s >= 50); result in *Eout.
__device__ int pade_block(int m, int aug, const double* al, const double* be, double theta,
                          double* sm, const double2** Eout)
{
    const int n = m + aug, nn = n * n;
    double2* X0 = reinterpret_cast<double2*>(sm);
    double2* X1 = reinterpret_cast<double2*>(sm + 2 * kNN);
    double* R0 = sm + 4 * kNN;
    double* R1 = R0 + nn;                 // R0, R1 contiguous: one complex n x n
    double* X0r = sm;
    double* X1r = sm + 2 * kNN;
    __shared__ int s_scale;
    const double cim = -theta;            // A = i cim Tbar  (propagate.cu:91-92)
    // infinity norm: row sums of |A_ij| in column order, as the oracle does,
    // one row per lane of wave 0, then the (exact) maximum over the rows
    if (threadIdx.x < 64) {
        const int i = threadIdx.x;
        double rs = 0.0;
        // the nonzeros of row i of Tbar in column order (tbar_at's pattern;
        // the zeros of the full row add +0 to a sum that is never -0, so
        // this is the full row's sum bit for bit)
        if (i < m) {
            if (i >= 1) rs += __builtin_fabs(cim * be[i - 1]);
            rs += __builtin_fabs(cim * al[i]);
            if (i + 1 < m) rs += __builtin_fabs(cim * be[i]);
            if (i == 0 && m < n) rs += __builtin_fabs(cim * 1.0);
        } else if (i < n) {
            if (i + 1 < n) rs += __builtin_fabs(cim * 1.0);
        }
        for (int off = 32; off > 0; off >>= 1) rs = __builtin_fmax(rs, __shfl_xor(rs, off, 64));
        if (threadIdx.x == 0) {
            double nrm = rs;
            int s = 0;
            while (nrm > 0.5) { nrm = 0.5 * nrm; s = s + 1; }
            s_scale = s;
        }
    }
    __syncthreads();
    const int s = s_scale;
    if (s >= 50) return -1;
    const double sc = __builtin_ldexp(1.0, -s);
    // c_k of padeClass.cu:161-168 for p = 6, same evaluation order
    double cc[7];
    cc[0] = 1;
    for (int k = 1; k <= 6; ++k) cc[k] = cc[k - 1] * (6.0 + 1. - k) / ((2 * 6.0 + 1. - k) * k);
    for (int e = threadIdx.x; e < nn; e += blockDim.x) {
        const int i = e / n, jj = e % n;
        R0[e] = (cim * tbar_at(i, jj, m, aug, al, be)) * sc;      // a
    }
    __syncthreads();
    mm_real(n, R0, R0, R1);            // P2
    mm_real(n, R1, R1, X0r);           // P4
    mm_real(n, R1, X0r, X0r + nn);     // P6
    for (int e = threadIdx.x; e < nn; e += blockDim.x) {
        const int i = e / n, jj = e % n;
        const double p2 = R1[e], p4 = X0r[e], p6 = X0r[nn + e];
        const double dI = (i == jj) ? 1.0 : 0.0;
        X1r[e] = cc[0] * dI - cc[2] * p2 + cc[4] * p4 - cc[6] * p6;   // V
        R1[e] = cc[1] * dI - cc[3] * p2 + cc[5] * p4;                 // W
    }
    __syncthreads();
    mm_real(n, R0, R1, X1r + nn);      // Ur = a W
    for (int e = threadIdx.x; e < nn; e += blockDim.x)
        X0[e] = make_double2(X1r[e], -X1r[nn + e]);                   // Den = V - i Ur
    __syncthreads();
    double2* Bm = reinterpret_cast<double2*>(R0);       // R0, R1 are dead
    for (int e = threadIdx.x; e < nn; e += blockDim.x) Bm[e] = make_double2(0.0, X1r[nn + e]); // Num = i Ur
    __syncthreads();
    // Gauss-Jordan with partial pivoting (|re|+|im|, first maximum) on
    // [Den | Num], from one pair of LDS buffers (G: Den part, Num part) into
    // the other (H) at every pivot: the row swap, the scaling of the pivot row
    // and the elimination are one pass and one barrier.  Every wave finds the
    // pivot itself from column k in its registers: it stays on the diagonal
    // unless a later row is strictly larger (one ballot), and only then a
    // shuffle reduction finds the first maximum; every thread takes the
    // pivot's reciprocal.  Each thread owns the same elements at every pivot:
    // their (row, column) pairs are computed once, outside the pivot loop.
    double2* Ga = X0;
    double2* Gb = Bm;
    double2* Ha = X1;
    double2* Hb = reinterpret_cast<double2*>(sm + 6 * kNN);
    constexpr int kPer = (2 * kNMax * kNMax + kPadeThreads - 1) / kPadeThreads;
    int er[kPer], ec[kPer];
#pragma unroll
    for (int r = 0; r < kPer; ++r) {
        const int e = threadIdx.x + r * kPadeThreads;
        er[r] = e < n * 2 * n ? e / (2 * n) : -1;
        ec[r] = e < n * 2 * n ? e % (2 * n) : 0;
    }
    const int lane = threadIdx.x & 63;
    for (int k = 0; k < n; ++k) {
        const int i0 = k + lane;
        double v = -1.0;
        int idx = n;
        double2 a0 = make_double2(0.0, 0.0);
        if (i0 < n) {
            a0 = Ga[i0 * n + k];
            v = __builtin_fabs(a0.x) + __builtin_fabs(a0.y);
            idx = i0;
        }
        int p = k;
        if (__ballot(v > readlane_f64(v, 0)) != 0) {
            for (int off = 32; off > 0; off >>= 1) {
                const double ov = __shfl_xor(v, off, 64);
                const int oi = __shfl_xor(idx, off, 64);
                if (ov > v || (ov == v && oi < idx)) { v = ov; idx = oi; }
            }
            p = __builtin_amdgcn_readfirstlane(idx);
        }
        const double px = readlane_f64(a0.x, p - k), py = readlane_f64(a0.y, p - k);
        const double den = px * px + py * py;
        const double ir = px / den, ii = -py / den;
#pragma unroll
        for (int r = 0; r < kPer; ++r) {
            const int i = er[r], jj = ec[r];
            if (i < 0) continue;
            const bool num = jj >= n;
            const int c = num ? jj - n : jj;
            if (!num && c < k) continue;                // eliminated columns are not read again
            const double2* G = num ? Gb : Ga;
            double2* H = num ? Hb : Ha;
            const double2 x = G[p * n + c];                // pivot row, scaled
            const double2 u = make_double2(x.x * ir - x.y * ii, x.x * ii + x.y * ir);
            if (i == k) {
                H[i * n + c] = u;
                continue;
            }
            const int si = i == p ? k : i;                 // row i after the swap of rows k and p
            const double2 f = Ga[si * n + k];
            double2 a = G[si * n + c];
            a.x -= f.x * u.x - f.y * u.y;
            a.y -= f.x * u.y + f.y * u.x;
            H[i * n + c] = a;
        }
        __syncthreads();
        double2* t0 = Ga; Ga = Ha; Ha = t0;
        double2* t1 = Gb; Gb = Hb; Hb = t1;
    }
    // E = I + 2X
    for (int e = threadIdx.x; e < nn; e += blockDim.x) {
        const int i = e / n, jj = e % n;
        double2 a = Gb[e];
        Gb[e] = make_double2(2.0 * a.x + (i == jj ? 1.0 : 0.0), 2.0 * a.y);
    }
    __syncthreads();
    double2* cur = Gb;
    double2* oth = Hb;
    for (int k = 0; k < s; ++k) {
        mm_cplx(n, cur, cur, oth);
        double2* t2 = cur; cur = oth; oth = t2;
    }
    *Eout = cur;
    return s;
}

__global__ __launch_bounds__(kPadeThreads) void k_pade(Dev d)
{
    extern __shared__ double sm[];
    __shared__ double al[kMaxM], be[kMaxM];
    __shared__ double s_hq;
    const int b = d.s0 + blockIdx.x;
    const int m = d.m, Bp = d.Bp;
    // |H v_m|^2 summed over the last SpMM's tiles and, batched, |v_m|^2 over
    // the last update's tiles (the plain path has it reduced in red[2m])
    double hq2 = 0.0, nu_b = 0.0;
    {
        // the first four waves, in the same order whatever the block size
        const double* ps = d.plain ? d.part : d.part_s;
        double s = 0.0, v = 0.0;
        if (threadIdx.x < kThreads) {
            for (int t = threadIdx.x; t < d.ntile_p; t += kThreads) s += ps[(size_t)t * Bp + b];
            if (!d.plain)
                for (int t = threadIdx.x; t < d.ntile_e; t += kThreads)
                    v += d.part[((size_t)(2 * m) * d.ntile_e + t) * Bp + b];
        }
        s = wave_sum(s);
        v = wave_sum(v);
        __shared__ double l4[2][4];
        if ((threadIdx.x & 63) == 0 && threadIdx.x < kThreads) { l4[0][threadIdx.x >> 6] = s; l4[1][threadIdx.x >> 6] = v; }
        __syncthreads();
        hq2 = (l4[0][0] + l4[0][1]) + (l4[0][2] + l4[0][3]);
        nu_b = (l4[1][0] + l4[1][1]) + (l4[1][2] + l4[1][3]);
    }
    if (threadIdx.x == 0) {
        // beta_{m-1} and s_m from the last update's Gram sums
        const double nu = d.plain ? d.red[(size_t)(2 * m) * Bp + b] : nu_b;
        double bet = __builtin_sqrt(nu);
        const double bprev = m >= 2 ? d.beta[(size_t)(m - 2) * Bp + b] : 0.0;
        const double aprev = d.alpha[(size_t)(m - 1) * Bp + b];
        int broke = d.broke[b];
        double sm_ = 0.0;
        if (broke || !(bet > d.tol * (__builtin_fabs(aprev) + bprev))) { broke = 1; bet = 0.0; }
        else sm_ = 1.0 / bet;
        d.broke[b] = broke;
        d.beta[(size_t)(m - 1) * Bp + b] = bet;
        d.sc[(size_t)m * Bp + b] = sm_;
        s_hq = sm_ * __builtin_sqrt(hq2);
    }
    __syncthreads();
    for (int k = threadIdx.x; k < m; k += kPadeThreads) {
        al[k] = d.alpha[(size_t)k * Bp + b];
        be[k] = d.beta[(size_t)k * Bp + b];
    }
    __syncthreads();
    if (b >= d.B) {
        for (int k = threadIdx.x; k < m; k += kPadeThreads) d.ycoef[(size_t)k * Bp + b] = make_double2(0.0, 0.0);
        return;
    }
    const double2* E;
    const int s = pade_block(m, d.aug, al, be, d.theta, sm, &E);
    if (s < 0) {
        if (threadIdx.x == 0) atomicOr(d.status, (int)kStatusScaling);
        for (int k = threadIdx.x; k < m; k += kPadeThreads) d.ycoef[(size_t)k * Bp + b] = make_double2(0.0, 0.0);
        return;
    }
    const int n = m + d.aug;
    // Lanczos: column 0 (propStep.cu:31-44).  Taylor-Krylov: |A^p psi| times
    // the rows 0..m-1 of column m + p - 1 (taylorKrylovClass.cu:372-397)
    const int ecol = d.tk ? n - 1 : 0;
    const double f = d.tk ? d.tk_nrm[b] : 1.0;
    for (int k = threadIdx.x; k < m; k += kPadeThreads) {
        const double2 y = E[k * n + ecol];
        const double sk = d.sc[(size_t)k * Bp + b] * f;
        d.ycoef[(size_t)k * Bp + b] = make_double2(y.x * sk, y.y * sk);
    }
    if (threadIdx.x == 0) {
        // Expokit estimate (padeClass.cu:749-782), clamped (systemClass.cu:2075-2077);
        // the Taylor-Krylov cycle adds |1 - |psi|| before the clamp (k_finish_norm)
        const double hm = __builtin_fabs(be[m - 1]);
        const double2 e1 = E[(m - 1) * n + m], e2 = E[(m - 1) * n + m + 1];
        const double err1 = hm * __builtin_sqrt(e1.x * e1.x + e1.y * e1.y);
        const double err2 = hm * s_hq * __builtin_sqrt(e2.x * e2.x + e2.y * e2.y);
        double err = err1 > err2 ? err2 / (1 - err2 / err1) : err1;
        if (!(err == err)) atomicOr(d.status, (int)kStatusNonfinite);
        if (d.tk) {
            d.tk_err[b] = err;
        } else {
            if (err < 1.110223e-16) err = 1.110223e-16;
            d.err_sum[b] += err;
            if (d.err_max) d.err_max[b] = __builtin_fmax(d.err_max[b], err);
        }
    }
}

// X0, X1, R0/R1 and the second Num buffer of the elimination: 8 kNN doubles
static size_t pade_lds_bytes() { return sizeof(double) * 8 * kNN; }
static hipError_t pade_lim();

hipError_t launch_pade(const Dev& d, hipStream_t s)
{
    if (hipError_t e = pade_lim(); e != hipSuccess) return e;
    hipLaunchKernelGGL(k_pade, dim3(d.sn), dim3(kPadeThreads), pade_lds_bytes(), s, d);
    return hipGetLastError();
}

__global__ __launch_bounds__(kPadeThreads) void k_pade_tbar(int m, const double* alpha, const double* beta,
                                                         double theta, double2* Eout, int* s_out)
{
    extern __shared__ double sm[];
    __shared__ double al[kMaxM], be[kMaxM];
    const int c = blockIdx.x;
    for (int k = threadIdx.x; k < m; k += kPadeThreads) {
        al[k] = alpha[(size_t)c * m + k];
        be[k] = beta[(size_t)c * m + k];
    }
    __syncthreads();
    const double2* E;
    const int s = pade_block(m, 2, al, be, theta, sm, &E);
    if (threadIdx.x == 0) s_out[c] = s;
    const int n = m + 2;
    if (s < 0) return;
    // column-major output
    for (int e = threadIdx.x; e < n * n; e += kPadeThreads) {
        const int i = e / n, jj = e % n;
        Eout[(size_t)c * n * n + (size_t)jj * n + i] = E[e];
    }
}

// lift the dynamic LDS limit of both Pade kernels (once per process)
static hipError_t pade_lim()
{
    static const hipError_t e = [] {
        hipError_t r = hipFuncSetAttribute(reinterpret_cast<const void*>(&k_pade),
                                           hipFuncAttributeMaxDynamicSharedMemorySize, (int)pade_lds_bytes());
        if (r == hipSuccess)
            r = hipFuncSetAttribute(reinterpret_cast<const void*>(&k_pade_tbar),
                                    hipFuncAttributeMaxDynamicSharedMemorySize, (int)pade_lds_bytes());
        return r;
    }();
    return e;
}

hipError_t launch_pade_tbar(int m, int count, const double* alpha, const double* beta,
                            double theta, double2* E, int* s_out, hipStream_t s)
{
    if (hipError_t e = pade_lim(); e != hipSuccess) return e;
    hipLaunchKernelGGL(k_pade_tbar, dim3(count), dim3(kPadeThreads), pade_lds_bytes(), s, m, alpha, beta,
                       theta, E, s_out);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Combine: psi[i][b] = y_0 psi[i][b] + sum_{k=1}^{m-1} (y_k s_k) v_k[i][b]
// (propStep.cu:31-44 as cublasZgemv; FP64 accumulation, psi kept in FP64
// for both precisions), partial |psi|^2 for the renormalisation.
// ---------------------------------------------------------------------------
template <bool PLAIN, typename VT, int MAXV>
__global__ __launch_bounds__(kThreads) void k_combine(Dev d)
{
    using V2 = typename Vec2<VT>::T;
    __shared__ double2 s_y[MAXV][64];
    Tile t = elem_tile<PLAIN>(d, d.rt_e, d.ntile_e);
    if (!t.ok) return;
    const int b = t.b, bl = PLAIN ? 0 : (threadIdx.x & 63);
    const int Bp = d.Bp, m = d.m;
    if (PLAIN) {
        if (threadIdx.x < (unsigned)m) s_y[threadIdx.x][0] = d.ycoef[threadIdx.x];
    } else if (threadIdx.x < 64) {
        for (int k = 0; k < m; ++k) s_y[k][bl] = d.ycoef[(size_t)k * Bp + b];
    }
    __syncthreads();
    double2 y[MAXV];
#pragma unroll
    for (int k = 0; k < MAXV; ++k) y[k] = (k < m) ? s_y[k][bl] : make_double2(0.0, 0.0);
    const size_t nB = (size_t)d.n * Bp;
    const V2* __restrict__ V = reinterpret_cast<const V2*>(d.V);
    double acc = 0.0;
    for (int i = t.rfirst; i < t.r1; i += t.rstep) {
        const size_t e = (size_t)i * Bp + b;
        const double2 p = d.psi[e];
        double ar = y[0].x * p.x - y[0].y * p.y;
        double ai = y[0].x * p.y + y[0].y * p.x;
#pragma unroll
        for (int k = 1; k < MAXV; ++k) {
            if (k < m) {
                const V2 v = V[(size_t)k * d.vs + e];
                const double xr = (double)v.x, xi = (double)v.y;
                ar += y[k].x * xr - y[k].y * xi;
                ai += y[k].x * xi + y[k].y * xr;
            }
        }
        d.psi[e] = make_double2(ar, ai);
        acc += ar * ar + ai * ai;
    }
    double out[1] = {acc};
    store_partials<PLAIN, 1>(out, 1, d.part, d.ntile_e, t.rtile, Bp, b);
}

template <bool PLAIN, typename VT>
static void launch_combine_t(const Dev& d, hipStream_t s)
{
    dim3 grid(elem_grid(d, d.ntile_e)), block(kThreads);
    if (d.m <= 8) hipLaunchKernelGGL((k_combine<PLAIN, VT, 8>), grid, block, 0, s, d);
    else if (d.m <= 16) hipLaunchKernelGGL((k_combine<PLAIN, VT, 16>), grid, block, 0, s, d);
    else if (d.m <= 24) hipLaunchKernelGGL((k_combine<PLAIN, VT, 24>), grid, block, 0, s, d);
    else hipLaunchKernelGGL((k_combine<PLAIN, VT, 32>), grid, block, 0, s, d);
}

hipError_t launch_combine(const Dev& d, hipStream_t s)
{
    if (!d.plain) return launch_combine_b(d, s);
    launch_combine_t<true, double>(d, s);
    return hipGetLastError();
}

template <bool PLAIN>
__global__ __launch_bounds__(kThreads) void k_finish_norm(Dev d, int advance_step)
{
    const int b = PLAIN ? 0 : d.s0 + blockIdx.x * 64 + (threadIdx.x & 63);
    const double s = block_tile_sum<PLAIN>(d.part, 0, d.ntile_e, d.Bp, b);
    if (threadIdx.x < (PLAIN ? 1u : 64u)) {
        d.nrm2[b] = s;
        if (d.tk && b < d.B) {
            // getTaylorKrylovError (taylorKrylovClass.cu:426-440): Krylov
            // estimate + |1 - |psi||, clamped (systemClass.cu:2075-2077)
            double err = d.tk_err[b] + __builtin_fabs(1. - __builtin_sqrt(s));
            if (err < 1.110223e-16) err = 1.110223e-16;
            d.err_sum[b] += err;
            if (d.err_max) d.err_max[b] = __builtin_fmax(d.err_max[b], err);
        }
    }
    if (advance_step && blockIdx.x == 0 && threadIdx.x == 0) *d.step = *d.step + 1;
}

hipError_t launch_finish_norm(const Dev& d, int advance_step, hipStream_t s)
{
    if (d.plain) hipLaunchKernelGGL((k_finish_norm<true>), dim3(1), dim3(kThreads), 0, s, d, advance_step);
    else hipLaunchKernelGGL((k_finish_norm<false>), dim3(d.sn / 64), dim3(kThreads), 0, s, d, advance_step);
    return hipGetLastError();
}

__global__ void k_advance_step(Dev d) { *d.step = *d.step + 1; }

hipError_t launch_advance_step(const Dev& d, hipStream_t s)
{
    hipLaunchKernelGGL(k_advance_step, dim3(1), dim3(1), 0, s, d);
    return hipGetLastError();
}

__global__ void k_set_unit_norm(Dev d)
{
    for (int b = threadIdx.x; b < d.Bp; b += blockDim.x) d.nrm2[b] = 1.0;
}

hipError_t launch_set_unit_norm(const Dev& d, hipStream_t s)
{
    hipLaunchKernelGGL(k_set_unit_norm, dim3(1), dim3(kThreads), 0, s, d);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Taylor-Krylov propagation (taylorKrylovClass.cu:217-434):
//   psi' = sum_{k<p} c^k A^k psi / k!  +  |A^p psi| Q_m [exp(c Tbar_p)]_{0:m, m+p-1}
// c = -i tau / hbar, A the noisy Hamiltonian of the trajectory, Q_m and
// Tbar_p (augmentation p) from the Lanczos decomposition of the constant
// auxiliary matrix on A^p psi / |A^p psi|.
//   tk_power(k), k = 1..p   u_k = A u_{k-1} (u_0 = psi); t += c^k u_k / k!
//                           for k < p (t = psi + c u_1 at k = 1); partial
//                           |u_p|^2 at k = p (batched: row-blocked, in
//                           dq_batched.hip)
//   tk_start                |u_p| per trajectory, basis slot 0 = u_p / |u_p|
//   Lanczos + pade on the auxiliary matrix (Dev.tk = 1, Dev.aug = p)
//   tk_combine              psi = t + sum_{k<m} y_k s_k v_k, partial |psi|^2
// The Taylor vectors are FP64 in both precisions: |A^k psi| grows as |A|^k.
// ---------------------------------------------------------------------------

// single trajectory: 16 lanes per row, as k_spmv_plain
template <typename VT>
__global__ __launch_bounds__(kThreads) void k_tk_power_plain(Dev d, int k)
{
    const int rtile = xcd_tile(blockIdx.x, gridDim.x);
    if (rtile >= d.ntile_s) return;
    const int tid = threadIdx.x, grp = tid >> 4, l = tid & 15;
    const size_t nB = (size_t)d.n;
    const double2* __restrict__ x = k == 1 ? d.psi : d.tk_u + (size_t)((k - 1) & 1) * nB;
    double2* __restrict__ y = d.tk_u + (size_t)(k & 1) * nB;
    const VT* __restrict__ diag = reinterpret_cast<const VT*>(d.diag);
    const bool last = k == d.tk_p;
    const double2 ck = last ? make_double2(0.0, 0.0) : d.tk_coef[k];
    const int nact = d.nact[0];
    double acc = 0.0;
    const int r0 = rtile * d.rt_s, r1 = min(d.n, r0 + d.rt_s);
    for (int i = r0 + grp; i < r1; i += 16) {
        const int p0 = d.rowptr[i], p1 = d.rowptr[i + 1];
        double sr = 0.0, si = 0.0;
        for (int q = p0 + l; q < p1; q += 16) {
            const double a = d.val64[q];
            const double2 xv = x[d.col[q]];
            sr += a * xv.x;
            si += a * xv.y;
        }
        sr = group_sum(sr, 16);
        si = group_sum(si, 16);
        if (l == 0) {
            const double2 xi = x[i];
            const double di = (double)diag[i];
            double wr = di * xi.x + d.g * sr;
            double wi = di * xi.y + d.g * si;
            if (i >= nact) { wr = 0.0; wi = 0.0; }
            y[i] = make_double2(wr, wi);
            if (!last) {
                double2 tv = k == 1 ? xi : d.tk_t[i];
                tv.x += ck.x * wr - ck.y * wi;
                tv.y += ck.x * wi + ck.y * wr;
                d.tk_t[i] = tv;
            } else {
                acc += wr * wr + wi * wi;
            }
        }
    }
    if (last) {
        double out[1] = {acc};
        store_partials<true, 1>(out, 1, d.part, d.ntile_s, rtile, 1, 0);
    }
}

hipError_t launch_tk_power(const Dev& d, int k, hipStream_t s)
{
    if (d.plain) {
        // the single-trajectory path is FP64 only
        hipLaunchKernelGGL((k_tk_power_plain<double>), dim3(pad8(d.ntile_s)), dim3(kThreads), 0, s, d, k);
    } else {
        return launch_tk_power_b(d, k, s);
    }
    return hipGetLastError();
}

// |u_p| per trajectory from the partials of tk_power(p), one per SpMM tile
template <bool PLAIN>
__global__ __launch_bounds__(kThreads) void k_tk_norm(Dev d)
{
    const int b = PLAIN ? 0 : d.s0 + blockIdx.x * 64 + (threadIdx.x & 63);
    const double s = block_tile_sum<PLAIN>(d.part, 0, d.ntile_s, d.Bp, b);
    if (threadIdx.x < (PLAIN ? 1u : 64u)) d.tk_nrm[b] = __builtin_sqrt(s);
}

// basis slot 0 = u_p / |u_p| (calcAndStoreNormAndScale of the auxiliary
// decomposition, taylorKrylovClass.cu:345)
template <bool PLAIN, typename VT>
__global__ __launch_bounds__(kThreads) void k_tk_scale(Dev d)
{
    using V2 = typename Vec2<VT>::T;
    Tile t = elem_tile<PLAIN>(d, d.rt_e, d.ntile_e);
    if (!t.ok) return;
    const int b = t.b;
    const double nu = d.tk_nrm[b];
    const double inv = nu > 0.0 ? 1.0 / nu : 0.0;
    const double2* __restrict__ u = d.tk_u + (size_t)(d.tk_p & 1) * ((size_t)d.n * d.Bp);
    V2* v0 = reinterpret_cast<V2*>(d.V);
    const int rend = PLAIN ? t.r1 : min(t.r1, (int)uniform(d.nact[b & ~63]));
    for (int i = t.rfirst; i < rend; i += t.rstep) {
        const size_t e = (size_t)i * d.Bp + b;
        const double2 a = u[e];
        V2 v;
        v.x = (VT)(a.x * inv);
        v.y = (VT)(a.y * inv);
        v0[e] = v;
    }
}

hipError_t launch_tk_start(const Dev& d, hipStream_t s)
{
    dim3 grid(elem_grid(d, d.ntile_e)), block(kThreads);
    if (d.plain) {
        hipLaunchKernelGGL((k_tk_norm<true>), dim3(1), block, 0, s, d);
        hipLaunchKernelGGL((k_tk_scale<true, double>), grid, block, 0, s, d);
    } else {
        hipLaunchKernelGGL((k_tk_norm<false>), dim3(d.sn / 64), block, 0, s, d);
        if (d.mixed) hipLaunchKernelGGL((k_tk_scale<false, float>), grid, block, 0, s, d);
        else hipLaunchKernelGGL((k_tk_scale<false, double>), grid, block, 0, s, d);
    }
    return hipGetLastError();
}

// psi = t + sum_{k<m} (y_k s_k) v_k, partial |psi|^2 (taylorKrylovClass.cu:372-418)
template <bool PLAIN, typename VT, int MAXV>
__global__ __launch_bounds__(kThreads) void k_tk_combine(Dev d)
{
    using V2 = typename Vec2<VT>::T;
    __shared__ double2 s_y[MAXV][64];
    Tile t = elem_tile<PLAIN>(d, d.rt_e, d.ntile_e);
    if (!t.ok) return;
    const int b = t.b, bl = PLAIN ? 0 : (threadIdx.x & 63);
    const int Bp = d.Bp, m = d.m;
    if (PLAIN) {
        if (threadIdx.x < (unsigned)m) s_y[threadIdx.x][0] = d.ycoef[threadIdx.x];
    } else if (threadIdx.x < 64) {
        for (int k = 0; k < m; ++k) s_y[k][bl] = d.ycoef[(size_t)k * Bp + b];
    }
    __syncthreads();
    double2 y[MAXV];
#pragma unroll
    for (int k = 0; k < MAXV; ++k) y[k] = (k < m) ? s_y[k][bl] : make_double2(0.0, 0.0);
    const size_t nB = (size_t)d.n * Bp;
    const V2* __restrict__ V = reinterpret_cast<const V2*>(d.V);
    const int rend = PLAIN ? t.r1 : min(t.r1, (int)uniform(d.nact[b & ~63]));
    double acc = 0.0;
    for (int i = t.rfirst; i < rend; i += t.rstep) {
        const size_t e = (size_t)i * Bp + b;
        const double2 tv = d.tk_t[e];
        double ar = 0.0, ai = 0.0;
#pragma unroll
        for (int k = 0; k < MAXV; ++k) {
            if (k < m) {
                const V2 v = V[(size_t)k * d.vs + e];
                const double xr = (double)v.x, xi = (double)v.y;
                ar += y[k].x * xr - y[k].y * xi;
                ai += y[k].x * xi + y[k].y * xr;
            }
        }
        ar = tv.x + ar;
        ai = tv.y + ai;
        d.psi[e] = make_double2(ar, ai);
        acc += ar * ar + ai * ai;
    }
    double out[1] = {acc};
    store_partials<PLAIN, 1>(out, 1, d.part, d.ntile_e, t.rtile, Bp, b);
}

template <bool PLAIN, typename VT>
static void launch_tk_combine_t(const Dev& d, hipStream_t s)
{
    dim3 grid(elem_grid(d, d.ntile_e)), block(kThreads);
    if (d.m <= 8) hipLaunchKernelGGL((k_tk_combine<PLAIN, VT, 8>), grid, block, 0, s, d);
    else if (d.m <= 16) hipLaunchKernelGGL((k_tk_combine<PLAIN, VT, 16>), grid, block, 0, s, d);
    else if (d.m <= 24) hipLaunchKernelGGL((k_tk_combine<PLAIN, VT, 24>), grid, block, 0, s, d);
    else hipLaunchKernelGGL((k_tk_combine<PLAIN, VT, 32>), grid, block, 0, s, d);
}

hipError_t launch_tk_combine(const Dev& d, hipStream_t s)
{
    if (!d.plain) return launch_combine_b(d, s);      // the batched combine with d.tk set
    launch_tk_combine_t<true, double>(d, s);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Spectra (spectraClass.cu:575-750): average C(n) over the batch, Gaussian
// window exp(-0.5((n+1)dt/sigma)^2) for n < T (gpuSpectraHelper.cu:16-42),
// zero-pad to nfft.  Then a radix-2 Stockham autosort FFT (replaces
// cufftExecZ2Z): one block per transform, ping-pong between two buffers.
// ---------------------------------------------------------------------------
// wts[Bp] (slot order, may be null): C(n) = sum_b wts[b] C_b(n) instead of
// the batch mean
__global__ void k_spectrum_prep(Dev d, int64_t T, double dt, double sigma, int nfft, double2* buf,
                                const double* wts)
{
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < nfft; k += gridDim.x * blockDim.x) {
        double2 v = make_double2(0.0, 0.0);
        if (k <= T) {
            double sr = 0.0, si = 0.0;
            for (int b = 0; b < d.Bp; ++b) {
                const double2 c = d.corr[(size_t)k * d.Bp + b];
                if (wts) {
                    sr += wts[b] * c.x;
                    si += wts[b] * c.y;
                } else if (b < d.B) {
                    sr += c.x;
                    si += c.y;
                }
            }
            double w = wts ? 1.0 : 1.0 / (double)d.B;
            if (sigma > 0.0 && k < T) {
                const double t = dt * (double)(k + 1);
                w *= exp(-0.5 * (t / sigma) * (t / sigma));
            }
            v = make_double2(sr * w, si * w);
        }
        buf[k] = v;
    }
}

hipError_t launch_spectrum_prep(const Dev& d, int64_t T, double dt, double sigma, int nfft,
                                double2* buf, const double* wts, hipStream_t s)
{
    hipLaunchKernelGGL(k_spectrum_prep, dim3((nfft + 255) / 256), dim3(256), 0, s, d, T, dt, sigma, nfft, buf, wts);
    return hipGetLastError();
}

__global__ __launch_bounds__(kThreads) void k_fft_stockham(int N, double2* a, double2* b, int sign)
{
    double2* x = a + (size_t)blockIdx.x * N;
    double2* y = b + (size_t)blockIdx.x * N;
    int n = N, s = 1;
    while (n > 1) {
        const int m = n >> 1;
        for (int t = threadIdx.x; t < (N >> 1); t += blockDim.x) {
            const int q = t % s, p = t / s;
            double sn, cs;
            sincospi(2.0 * (double)p / (double)n, &sn, &cs);
            const double wr = cs, wi = sign >= 0 ? sn : -sn;
            const double2 u = x[q + s * p];
            const double2 v = x[q + s * (p + m)];
            y[q + s * (2 * p)] = make_double2(u.x + v.x, u.y + v.y);
            const double dr = u.x - v.x, di = u.y - v.y;
            y[q + s * (2 * p + 1)] = make_double2(dr * wr - di * wi, dr * wi + di * wr);
        }
        __syncthreads();
        double2* tmp = x; x = y; y = tmp;
        n = m;
        s <<= 1;
    }
    // result is in x; make sure it ends in buffer a
    if (x != a + (size_t)blockIdx.x * N) {
        for (int t = threadIdx.x; t < N; t += blockDim.x) y[t] = x[t];
    }
}

hipError_t launch_fft(int n, int batch, double2* a, double2* b, int sign, hipStream_t s)
{
    hipLaunchKernelGGL(k_fft_stockham, dim3(batch), dim3(kThreads), 0, s, n, a, b, sign);
    return hipGetLastError();
}

}  // namespace dq
