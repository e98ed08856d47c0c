// dq_batched.hip -- batched (B >= 64 padded) kernels of the Krylov cycle.
//
// SpMM: one wave per block of 16 rows x 64 trajectories over the union of
// the rows' columns (register reuse of every gathered x run).
//
// Update and combine: a wave covers 16 trajectories x 4 groups of Krylov
// vectors (lane = 16 * kgroup + t).  Each lane streams only the vectors of
// its group, the four partial sums of a row are combined with two xor
// shuffles (no barrier), and each lane keeps FP64 Gram accumulators only
// for its own group's vectors -- a quarter of the registers of a
// one-lane-per-trajectory layout, so several waves per SIMD stay resident.
#include "dq_device.h"
#include "dq_internal.h"

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
// SpMM  w = d .* x + g * J x  (lanczosClass.cu:618-661 per trajectory).
// Row-blocked: a wave owns kRB consecutive rows x 64 trajectories and walks
// the union of their columns, so each gathered x run (64 trajectories, one
// contiguous 512 B / 1 KB segment) feeds kRB rows from registers.  For the
// dipole tubes kRB = 16 cuts the gathers per row from ~61 to ~10.  Column
// indices and the kRB values per column are fetched lane-parallel and
// broadcast with readlane.
// ---------------------------------------------------------------------------
template <typename VT>
__global__ __launch_bounds__(kThreads) void k_spmm_b(Dev d, int j, int last)
{
    using V2 = typename BVec2<VT>::T;
    __shared__ VT s_ar[2][kRB][64], s_ai[2][kRB][64];
    const int ntt = d.Bp / 64;
    const int logical = xcd_tile(blockIdx.x, gridDim.x);
    if (logical >= d.ntile_s * ntt) return;
    const int rb = logical / ntt, ttile = logical % ntt;     // one row block per block
    const int lane = threadIdx.x & 63;
    const int wave = uniform(threadIdx.x >> 6);
    const int b = ttile * 64 + lane;
    const int Bp = d.Bp;
    const size_t nB = (size_t)d.n * Bp;
    const V2* __restrict__ x = reinterpret_cast<const V2*>(d.V) + (size_t)j * nB;
    V2* __restrict__ y = reinterpret_cast<V2*>(d.V) + (size_t)(j + 1) * nB;
    const VT* __restrict__ diag = reinterpret_cast<const VT*>(d.diag);
    const VT* __restrict__ rv;
    if constexpr (sizeof(VT) == 4) rv = d.rb_val32;
    else rv = d.rb_val64;
    // the four waves take contiguous quarters (multiples of 8) of the union
    const int u0 = uniform(d.rb_ptr[rb]), u1 = uniform(d.rb_ptr[rb + 1]);
    const int len = u1 - u0;
    const int quarter = ((len + 63) / 64) * 16;
    const int w0 = min(u1, u0 + wave * quarter), w1 = min(u1, w0 + quarter);
    VT ar[kRB], ai[kRB];
#pragma unroll
    for (int r = 0; r < kRB; ++r) { ar[r] = 0; ai[r] = 0; }
    for (int cb = w0; cb < w1; cb += 64) {
        const int lim = min(64, w1 - cb);
        const int mc = lane < lim ? d.rb_col[cb + lane] : 0;
        for (int q = 0; q < lim; q += 16) {
            // 16 gathers and the 16 x kRB values of these columns in flight
            const size_t vb = (size_t)(cb + q) * kRB;
            const size_t vend = (size_t)w1 * kRB;
            VT mv[4];
#pragma unroll
            for (int h = 0; h < 4; ++h) mv[h] = vb + 64 * h + lane < vend ? rv[vb + 64 * h + lane] : (VT)0;
            V2 xv[16];
#pragma unroll
            for (int u = 0; u < 16; ++u) {
                if (q + u < lim) {
                    const int c = __builtin_amdgcn_readlane(mc, q + u);
                    xv[u] = x[(size_t)c * Bp + b];
                }
            }
#pragma unroll
            for (int u = 0; u < 16; ++u) {
                if (q + u < lim) {
#pragma unroll
                    for (int r = 0; r < kRB; ++r) {
                        const VT v = bcast(mv[u >> 2], (u & 3) * kRB + r);
                        ar[r] += v * xv[u].x;
                        ai[r] += v * xv[u].y;
                    }
                }
            }
        }
    }
    // deterministic tree over the waves: (w0 + w2) + (w1 + w3)
    if (wave >= 2) {
#pragma unroll
        for (int r = 0; r < kRB; ++r) { s_ar[wave - 2][r][lane] = ar[r]; s_ai[wave - 2][r][lane] = ai[r]; }
    }
    __syncthreads();
    if (wave < 2) {
#pragma unroll
        for (int r = 0; r < kRB; ++r) { ar[r] += s_ar[wave][r][lane]; ai[r] += s_ai[wave][r][lane]; }
    }
    __syncthreads();
    if (wave == 1) {
#pragma unroll
        for (int r = 0; r < kRB; ++r) { s_ar[0][r][lane] = ar[r]; s_ai[0][r][lane] = ai[r]; }
    }
    __syncthreads();
    if (wave == 0) {
#pragma unroll
        for (int r = 0; r < kRB; ++r) { s_ar[1][r][lane] = ar[r] + s_ar[0][r][lane]; s_ai[1][r][lane] = ai[r] + s_ai[0][r][lane]; }
    }
    __syncthreads();
    // epilogue: wave w finishes rows w, w+4, ... of the block
    const int nact = d.nact[b];
    const VT g = (VT)d.g;
    double acc0 = 0.0, acc1 = 0.0;
#pragma unroll
    for (int rr = 0; rr < kRB / 4; ++rr) {
        const int r = wave + 4 * rr;
        const int i = rb * kRB + r;
        if (i < d.n) {
            const VT sr = s_ar[1][r][lane];
            const VT si = s_ai[1][r][lane];
            const size_t e = (size_t)i * Bp + b;
            const V2 xi = x[e];
            const VT di = diag[e];
            VT wr = di * xi.x + g * sr;
            VT wi = di * xi.y + g * si;
            if (i >= nact) { wr = 0; wi = 0; }
            if (!last) {
                V2 w;
                w.x = wr;
                w.y = wi;
                y[e] = w;
                const double xr = (double)xi.x, xm = (double)xi.y, dr = (double)wr, dm = (double)wi;
                acc0 += xr * dr + xm * dm;
                acc1 += xr * dm - xm * dr;
            } else {
                acc0 += (double)wr * (double)wr + (double)wi * (double)wi;
            }
        }
    }
    double out[2] = {acc0, acc1};
    bstore_partials<2>(out, last ? 1 : 2, d.part, d.ntile_s, rb, Bp, b);
}

hipError_t launch_spmm_b(const Dev& d, int j, int last, hipStream_t s)
{
    dim3 grid(bpad8(d.ntile_s * (d.Bp / 64)));
    if (d.mixed) hipLaunchKernelGGL((k_spmm_b<float>), grid, dim3(kThreads), 0, s, d, j, last);
    else hipLaunchKernelGGL((k_spmm_b<double>), grid, dim3(kThreads), 0, s, d, j, last);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Per-trajectory scalars of iteration j (batched): grid (Bp/64, ceil(j/4) or 1).
// Every block sums <v_j|w> over the SpMM tiles and rebuilds s_j and the Gram
// row for its 64 trajectories; block y == 0 stores them; wave w of block y
// computes the projection h_k, k = 4y + w < j.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kThreads) void k_scalars_hb(Dev d, int j)
{
    __shared__ double s_sj[64];
    __shared__ double2 s_G[kMaxM + 1][64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int b = blockIdx.x * 64 + lane;
    const int Bp = d.Bp;
    // <v_j|w> reduced over the SpMM tiles by k_reduce_wide into red[kRedC..]
    const double cr = d.red[((size_t)kRedC(d.m)) * Bp + b];
    const double ci = d.red[((size_t)kRedC(d.m) + 1) * Bp + b];
    const bool writer = blockIdx.y == 0;
    if (wave == 0) {
        double sj = 1.0;
        if (j == 0) {
            s_G[0][lane] = make_double2(1.0, 0.0);
            if (writer) { d.sc[b] = 1.0; d.broke[b] = 0; }
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
            if (writer) {
                d.broke[b] = broke;
                d.beta[(size_t)(j - 1) * Bp + b] = bet;
                d.sc[(size_t)j * Bp + b] = sj;
            }
            for (int l = 0; l < j; ++l) {
                const double f = d.sc[(size_t)l * Bp + b] * sj;
                const double2 G = make_double2(f * d.red[(size_t)(2 * l) * Bp + b],
                                               f * d.red[(size_t)(2 * l + 1) * Bp + b]);
                s_G[l][lane] = G;
                if (writer) d.gram[(size_t)l * Bp + b] = G;
            }
            s_G[j][lane] = make_double2(1.0, 0.0);
            s_sj[lane] = bet;          // beta_{j-1}, needed by k = j-1 below
        }
        if (writer) {
            const double hr = sj * sj * cr, hi = sj * sj * ci;
            d.alpha[(size_t)j * Bp + b] = hr;
            d.hco[((size_t)j * kMaxM + j) * Bp + b] = make_double2(hr, hi);
        }
    }
    __syncthreads();
    const int k = blockIdx.y * 4 + wave;
    if (k < j) {
        const double bk = (k == j - 1) ? s_sj[lane] : d.beta[(size_t)k * Bp + b];
        const double2 Gk1 = s_G[k + 1][lane];
        double sr = bk * Gk1.x, si = bk * Gk1.y;
        for (int l = 0; l <= k; ++l) {
            const double2 a = d.hco[((size_t)k * kMaxM + l) * Bp + b];
            const double2 G = s_G[l][lane];
            sr += a.x * G.x + a.y * G.y;
            si += a.x * G.y - a.y * G.x;
        }
        d.hco[((size_t)j * kMaxM + k) * Bp + b] = make_double2(sr, si);
    }
}

// red[dst + q][b] = sum over tiles of part[q][tile][b]: 16 waves per block,
// four load chains per lane, fixed combination order
__global__ __launch_bounds__(1024) void k_reduce_wide(Dev d, int ntile, int dst)
{
    __shared__ double lds[16][64];
    const int q = blockIdx.y;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int b = blockIdx.x * 64 + lane;
    const int Bp = d.Bp;
    double a[4] = {0.0, 0.0, 0.0, 0.0};
    int t = wave;
    for (; t + 48 < ntile; t += 64) {
#pragma unroll
        for (int u = 0; u < 4; ++u) a[u] += d.part[((size_t)q * ntile + t + 16 * u) * Bp + b];
    }
    for (; t < ntile; t += 16) a[0] += d.part[((size_t)q * ntile + t) * Bp + b];
    lds[wave][lane] = (a[0] + a[1]) + (a[2] + a[3]);
    __syncthreads();
    if (wave == 0) {
        double s = 0.0;
        for (int w = 0; w < 16; ++w) s += lds[w][lane];
        d.red[((size_t)dst + q) * Bp + b] = s;
    }
}

hipError_t launch_reduce_wide(const Dev& d, int nq, int ntile, int dst, hipStream_t s)
{
    hipLaunchKernelGGL(k_reduce_wide, dim3(d.Bp / 64, nq), dim3(1024), 0, s, d, ntile, dst);
    return hipGetLastError();
}

hipError_t launch_scalars_hb(const Dev& d, int j, hipStream_t s)
{
    hipError_t e = launch_reduce_wide(d, 2, d.ntile_s, kRedC(d.m), s);
    if (e != hipSuccess) return e;
    const int ny = j > 0 ? (j + 3) / 4 : 1;
    hipLaunchKernelGGL(k_scalars_hb, dim3(d.Bp / 64, ny), dim3(kThreads), 0, s, d, j);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Update of iteration j: v_{j+1} = s_j w - sum_{k<=j} (h_k s_k) v_k and the
// FP64 partial Gram row <v_k|v_{j+1}>, |v_{j+1}|^2.  Block = rt_e rows x 16
// trajectories; lane = 16 kg + t handles vectors k = kg, kg+4, ...
// ---------------------------------------------------------------------------
template <typename VT, int KW>
__global__ __launch_bounds__(kThreads) void k_update_b(Dev d, int j)
{
    using V2 = typename BVec2<VT>::T;
    __shared__ double2 s_c[4 * KW][16];
    __shared__ double s_w[16];
    __shared__ double lds[4][2 * KW + 1][64];
    const int ntt = d.Bp / 16;
    const int logical = xcd_tile(blockIdx.x, gridDim.x);
    if (logical >= d.ntile_e * ntt) return;
    const int rtile = logical / ntt, ttile = logical % ntt;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int t = lane & 15, kg = lane >> 4;
    const int b = ttile * 16 + t;
    const int Bp = d.Bp;
    const size_t nB = (size_t)d.n * Bp;
    for (int q = threadIdx.x; q < 16 * (j + 1); q += kThreads) {
        const int k = q >> 4, tt = q & 15, bb = ttile * 16 + tt;
        const double2 h = d.hco[((size_t)j * kMaxM + k) * Bp + bb];
        const double sk = d.sc[(size_t)k * Bp + bb];
        s_c[k][tt] = make_double2(-h.x * sk, -h.y * sk);
        if (k == 0) s_w[tt] = d.sc[(size_t)j * Bp + bb];
    }
    __syncthreads();
    const V2* __restrict__ V = reinterpret_cast<const V2*>(d.V);
    V2* __restrict__ out = reinterpret_cast<V2*>(d.V) + (size_t)(j + 1) * nB;
    const double sw = s_w[t];
    double2 cf[KW];
#pragma unroll
    for (int u = 0; u < KW; ++u) {
        const int k = kg + 4 * u;
        cf[u] = (k <= j) ? s_c[k][t] : make_double2(0.0, 0.0);
    }
    double acc[2 * KW + 1];
#pragma unroll
    for (int q = 0; q < 2 * KW + 1; ++q) acc[q] = 0.0;
    const int r0 = rtile * d.rt_e, r1 = min(d.n, r0 + d.rt_e);
    // software pipeline: the vectors of row i+4 load while row i is combined
    V2 nk[KW], nw;
    if (r0 + wave < r1) {
        const size_t e = (size_t)(r0 + wave) * Bp + b;
#pragma unroll
        for (int u = 0; u < KW; ++u) {
            const int k = kg + 4 * u;
            if (k <= j) nk[u] = V[(size_t)k * nB + e];
        }
        if (kg == 0) nw = out[e];
    }
    for (int i = r0 + wave; i < r1; i += 4) {
        const size_t e = (size_t)i * Bp + b;
        V2 vk[KW];
#pragma unroll
        for (int u = 0; u < KW; ++u) vk[u] = nk[u];
        const V2 w = nw;
        if (i + 4 < r1) {
            const size_t en = e + (size_t)4 * Bp;
#pragma unroll
            for (int u = 0; u < KW; ++u) {
                const int k = kg + 4 * u;
                if (k <= j) nk[u] = V[(size_t)k * nB + en];
            }
            if (kg == 0) nw = out[en];
        }
        double pr = 0.0, pi = 0.0;
        if (kg == 0) {
            pr = sw * (double)w.x;
            pi = sw * (double)w.y;
        }
#pragma unroll
        for (int u = 0; u < KW; ++u) {
            const int k = kg + 4 * u;
            if (k <= j) {
                const double xr = (double)vk[u].x, xi = (double)vk[u].y;
                pr += cf[u].x * xr - cf[u].y * xi;
                pi += cf[u].x * xi + cf[u].y * xr;
            }
        }
        pr += __shfl_xor(pr, 16, 64);
        pi += __shfl_xor(pi, 16, 64);
        pr += __shfl_xor(pr, 32, 64);
        pi += __shfl_xor(pi, 32, 64);
        V2 nv;
        nv.x = (VT)pr;
        nv.y = (VT)pi;
        if (kg == 0) out[e] = nv;
        const double nr = (double)nv.x, ni = (double)nv.y;
#pragma unroll
        for (int u = 0; u < KW; ++u) {
            const int k = kg + 4 * u;
            if (k <= j) {
                const double xr = (double)vk[u].x, xi = (double)vk[u].y;
                acc[2 * u] += xr * nr + xi * ni;
                acc[2 * u + 1] += xr * ni - xi * nr;
            }
        }
        if (kg == 0) acc[2 * KW] += nr * nr + ni * ni;
    }
    // combine the four waves (rows) per lane, then write q = 2k, 2k+1 and nu
#pragma unroll
    for (int q = 0; q < 2 * KW + 1; ++q) lds[wave][q][lane] = acc[q];
    __syncthreads();
    if (wave == 0) {
#pragma unroll
        for (int u = 0; u < KW; ++u) {
            const int k = kg + 4 * u;
            if (k <= j) {
                for (int c = 0; c < 2; ++c) {
                    const int q = 2 * u + c;
                    const double v = (lds[0][q][lane] + lds[1][q][lane]) + (lds[2][q][lane] + lds[3][q][lane]);
                    d.part[((size_t)(2 * k + c) * d.ntile_e + rtile) * Bp + b] = v;
                }
            }
        }
        if (kg == 0) {
            const int q = 2 * KW;
            const double v = (lds[0][q][lane] + lds[1][q][lane]) + (lds[2][q][lane] + lds[3][q][lane]);
            d.part[((size_t)(2 * (j + 1)) * d.ntile_e + rtile) * Bp + b] = v;
        }
    }
}

hipError_t launch_update_b(const Dev& d, int j, hipStream_t s)
{
    dim3 grid(bpad8(d.ntile_e * (d.Bp / 16))), block(kThreads);
    const int nv = j + 1;
    if (d.mixed) {
        if (nv <= 8) hipLaunchKernelGGL((k_update_b<float, 2>), grid, block, 0, s, d, j);
        else if (nv <= 16) hipLaunchKernelGGL((k_update_b<float, 4>), grid, block, 0, s, d, j);
        else if (nv <= 24) hipLaunchKernelGGL((k_update_b<float, 6>), grid, block, 0, s, d, j);
        else hipLaunchKernelGGL((k_update_b<float, 8>), grid, block, 0, s, d, j);
    } else {
        if (nv <= 8) hipLaunchKernelGGL((k_update_b<double, 2>), grid, block, 0, s, d, j);
        else if (nv <= 16) hipLaunchKernelGGL((k_update_b<double, 4>), grid, block, 0, s, d, j);
        else if (nv <= 24) hipLaunchKernelGGL((k_update_b<double, 6>), grid, block, 0, s, d, j);
        else hipLaunchKernelGGL((k_update_b<double, 8>), grid, block, 0, s, d, j);
    }
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Combine psi = y_0 psi + sum_{k>=1} (y_k s_k) v_k (propStep.cu:31-44) and the
// partial |psi|^2; same 16 x 4 lane layout as the update.
// ---------------------------------------------------------------------------
template <typename VT, int KW>
__global__ __launch_bounds__(kThreads) void k_combine_b(Dev d)
{
    using V2 = typename BVec2<VT>::T;
    __shared__ double2 s_y[4 * KW][16];
    __shared__ double lds[4][64];
    const int ntt = d.Bp / 16;
    const int logical = xcd_tile(blockIdx.x, gridDim.x);
    if (logical >= d.ntile_e * ntt) return;
    const int rtile = logical / ntt, ttile = logical % ntt;
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
    const int r0 = rtile * d.rt_e, r1 = min(d.n, r0 + d.rt_e);
    V2 nk[KW];
    double2 np = make_double2(0.0, 0.0);
    if (r0 + wave < r1) {
        const size_t e = (size_t)(r0 + wave) * Bp + b;
#pragma unroll
        for (int u = 0; u < KW; ++u) {
            const int k = kg + 4 * u;
            if (k >= 1 && k < m) nk[u] = V[(size_t)k * nB + e];
        }
        if (kg == 0) np = d.psi[e];
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
                if (k >= 1 && k < m) nk[u] = V[(size_t)k * nB + en];
            }
            if (kg == 0) np = d.psi[en];
        }
        double pr = 0.0, pi = 0.0;
        if (kg == 0) {
            pr = y[0].x * p.x - y[0].y * p.y;
            pi = y[0].x * p.y + y[0].y * p.x;
        }
#pragma unroll
        for (int u = 0; u < KW; ++u) {
            const int k = kg + 4 * u;
            if (k >= 1 && k < m) {
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

hipError_t launch_combine_b(const Dev& d, hipStream_t s)
{
    dim3 grid(bpad8(d.ntile_e * (d.Bp / 16))), block(kThreads);
    if (d.mixed) {
        if (d.m <= 8) hipLaunchKernelGGL((k_combine_b<float, 2>), grid, block, 0, s, d);
        else if (d.m <= 16) hipLaunchKernelGGL((k_combine_b<float, 4>), grid, block, 0, s, d);
        else if (d.m <= 24) hipLaunchKernelGGL((k_combine_b<float, 6>), grid, block, 0, s, d);
        else hipLaunchKernelGGL((k_combine_b<float, 8>), grid, block, 0, s, d);
    } else {
        if (d.m <= 8) hipLaunchKernelGGL((k_combine_b<double, 2>), grid, block, 0, s, d);
        else if (d.m <= 16) hipLaunchKernelGGL((k_combine_b<double, 4>), grid, block, 0, s, d);
        else if (d.m <= 24) hipLaunchKernelGGL((k_combine_b<double, 6>), grid, block, 0, s, d);
        else hipLaunchKernelGGL((k_combine_b<double, 8>), grid, block, 0, s, d);
    }
    return hipGetLastError();
}

}  // namespace dq
