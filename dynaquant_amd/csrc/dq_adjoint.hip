// dq_adjoint.hip -- reverse sweep of the batched FP64 propagation
// (gradient of a trajectory loss with respect to the site energies and the
// coupling scale).  The reference has no derivatives; this differentiates
// the discrete map the engine computes:
//
//   per Krylov cycle (input q_0 = psi, normalised output psi''):
//     psi'' = psi'/|psi'|                    renorm reverse   (k_adj_renorm_*)
//     psi'  = sum_k y_k q_k                  ybar_k = q_k^H psibar', qbar_k += conj(y_k) psibar'
//     y     = exp(-i theta Tbar)[:, 0]       Pade reverse in one block per trajectory (k_adj_pade)
//     Lanczos, reversed as the three-term recurrence on the stored basis
//       (w' = H q_j - alpha_j q_j - beta_{j-1} q_{j-1}; equal to the engine's
//       full projection in exact arithmetic, so the derivative agrees to the
//       basis' orthogonality; O(m) vector passes per cycle):
//       wbar' = qbar_{j+1}/beta_j + (betabar_j - Re<qbar_{j+1}|q_{j+1}>/beta_j) q_{j+1}
//       alphabar = alphabar_T - Re<q_j|wbar'>, betabar_{j-1} -= Re<q_{j-1}|wbar'>
//       qbar_j += -alpha_j wbar' + alphabar (H q_j) + H wbar,  qbar_{j-1} -= beta_{j-1} wbar'
//       wbar = wbar' + alphabar q_j,  dbar_i += Re(conj(wbar_i) q_j,i),  gbar += Re<J wbar|q_j>
//
// Vectors are [i][b] double2 like the forward; per-trajectory scalars [k][Bp].
#include "dq_device.h"
#include "dq_internal.h"
#include "dq_adjoint.h"

namespace dq {

static inline int pad8a(int x) { return (x + 7) & ~7; }

// partial sums of the block (lane = trajectory, 4 waves summed in LDS)
template <int NQ>
__device__ __forceinline__ void adj_store_partials(const double (&acc)[NQ], int nq, double* part,
                                                   int ntile, int tile, int Bp, int b)
{
    __shared__ double lds[4][64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (int q = 0; q < NQ; ++q) {
        if (q >= nq) break;
        __syncthreads();
        lds[wave][lane] = acc[q];
        __syncthreads();
        if (wave == 0)
            part[((size_t)q * ntile + tile) * Bp + b] = (lds[0][lane] + lds[1][lane]) + (lds[2][lane] + lds[3][lane]);
    }
}

struct ATile { int rtile, b, r1, rfirst; bool ok; };

__device__ __forceinline__ ATile adj_tile(const Dev& d, int rt, int ntile)
{
    ATile t;
    const int ntt = d.Bp / 64;
    const int logical = xcd_tile(blockIdx.x, gridDim.x);
    t.ok = logical < ntile * ntt;
    t.rtile = logical / ntt;
    t.b = (logical % ntt) * 64 + (threadIdx.x & 63);
    const int r0 = t.rtile * rt;
    t.r1 = min(d.n, r0 + rt);
    t.rfirst = r0 + (threadIdx.x >> 6);
    return t;
}

__device__ __forceinline__ double2 cmul(double2 a, double2 b) { return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x); }

// ---------------------------------------------------------------------------
// renormalisation reverse, part 1: partial Re<psi''|pbar>
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kThreads) void k_adj_renorm_dot(Dev d, AdjDev a, int c)
{
    ATile t = adj_tile(d, d.rt_e, d.ntile_e);
    if (!t.ok) return;
    const size_t nB = (size_t)d.n * d.Bp;
    const double2* out = a.psi_out(c);
    double acc[1] = {0.0};
    for (int i = t.rfirst; i < t.r1; i += 4) {
        const size_t e = (size_t)i * d.Bp + t.b;
        const double2 p = out[e], q = a.pbar[e];
        acc[0] += p.x * q.x + p.y * q.y;
    }
    (void)nB;
    adj_store_partials<1>(acc, 1, d.part, d.ntile_e, t.rtile, d.Bp, t.b);
}

// part 2: psibar' = (pbar - psi'' r)/|psi'|, partial ybar_k = <q_k|psibar'>.
// Lane = 16 kg + t: 16 trajectories per workgroup, lane group kg owns the
// basis vectors k = kg + 4u, so each lane keeps 2 KW partial sums and no
// shuffles are needed; the four waves (rows) meet in LDS at the end.
template <int KW>
__global__ __launch_bounds__(kThreads) void k_adj_renorm_apply(Dev d, AdjDev a, int c)
{
    __shared__ double lds[4][2 * KW][64];
    const int ntt = d.Bp / 16;
    const int logical = xcd_tile(blockIdx.x, gridDim.x);
    if (logical >= d.ntile_e * ntt) return;
    const int rtile = logical / ntt, ttile = logical % ntt;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int t = lane & 15, kg = lane >> 4;
    const int b = ttile * 16 + t;
    const int Bp = d.Bp, m = d.m;
    const size_t nB = (size_t)d.n * Bp;
    const double r = d.red[b];
    const double inv = 1.0 / __builtin_sqrt(a.nrm2[(size_t)c * Bp + b]);
    const double2* out = a.psi_out(c);
    const double2* V = a.basis(c);
    const double* sc = a.sc + (size_t)c * (m + 1) * Bp;
    double acc[2 * KW];
#pragma unroll
    for (int q = 0; q < 2 * KW; ++q) acc[q] = 0.0;
    const int r0 = rtile * d.rt_e, r1 = min(d.n, r0 + d.rt_e);
    for (int i = r0 + wave; i < r1; i += 4) {
        const size_t e = (size_t)i * Bp + b;
        const double2 p = out[e], pb = a.pbar[e];
        const double2 v = make_double2((pb.x - p.x * r) * inv, (pb.y - p.y * r) * inv);
        if (kg == 0) a.pbar1[e] = v;
#pragma unroll
        for (int u = 0; u < KW; ++u) {
            const int k = kg + 4 * u;
            if (k < m) {
                const double2 q = V[(size_t)k * a.vs + e];
                acc[2 * u] += q.x * v.x + q.y * v.y;
                acc[2 * u + 1] += q.x * v.y - q.y * v.x;
            }
        }
    }
#pragma unroll
    for (int q = 0; q < 2 * KW; ++q) lds[wave][q][lane] = acc[q];
    __syncthreads();
    if (wave == 0) {
#pragma unroll
        for (int u = 0; u < KW; ++u) {
            const int k = kg + 4 * u;
            if (k >= m) continue;
            const double sk = sc[(size_t)k * Bp + b];        // q_k = s_k v_k
            for (int cc = 0; cc < 2; ++cc) {
                const int q = 2 * u + cc;
                const double v = (lds[0][q][lane] + lds[1][q][lane]) + (lds[2][q][lane] + lds[3][q][lane]);
                d.part[((size_t)(2 * k + cc) * d.ntile_e + rtile) * Bp + b] = sk * v;
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Pade reverse (padeClass.cu:204-706 forward), one block per trajectory.
// General complex restatement with the forward intermediates in a global
// scratch area; products and solves staged through LDS.
// ---------------------------------------------------------------------------
constexpr int kN = kMaxM + 2;
constexpr int kNNa = kN * kN;

// C (+)= op(A) op(B): ta/tb = 0 plain, 1 conjugate transpose
__device__ void gmm(int n, const double2* A, int ta, const double2* B, int tb, double2* C, int accumulate,
                    double2* lA, double2* lB)
{
    const int nn = n * n;
    for (int e = threadIdx.x; e < nn; e += blockDim.x) {
        const int i = e / n, j = e % n;
        double2 av = ta ? A[j * n + i] : A[e];
        if (ta) av.y = -av.y;
        lA[e] = av;
        double2 bv = tb ? B[j * n + i] : B[e];
        if (tb) bv.y = -bv.y;
        lB[e] = bv;
    }
    __syncthreads();
    for (int e = threadIdx.x; e < nn; e += blockDim.x) {
        const int i = e / n, j = e % n;
        double sr = 0.0, si = 0.0;
        for (int k = 0; k < n; ++k) {
            const double2 x = lA[i * n + k], y = lB[k * n + j];
            sr += x.x * y.x - x.y * y.y;
            si += x.x * y.y + x.y * y.x;
        }
        if (accumulate) { double2 o = C[e]; C[e] = make_double2(o.x + sr, o.y + si); }
        else C[e] = make_double2(sr, si);
    }
    __syncthreads();
}

// solve op(M) X = B in place in B (Gauss-Jordan, partial pivoting), op = conj-transpose if tm
__device__ void gsolve(int n, const double2* M, int tm, double2* B, double2* lM, double2* lB)
{
    __shared__ int s_piv;
    __shared__ double2 s_f[kN];
    const int nn = n * n;
    for (int e = threadIdx.x; e < nn; e += blockDim.x) {
        const int i = e / n, j = e % n;
        double2 v = tm ? M[j * n + i] : M[e];
        if (tm) v.y = -v.y;
        lM[e] = v;
        lB[e] = B[e];
    }
    __syncthreads();
    for (int k = 0; k < n; ++k) {
        if (threadIdx.x < 64) {
            const int i = k + threadIdx.x;
            double v = -1.0;
            int idx = n;
            if (i < n) { const double2 x = lM[i * n + k]; v = __builtin_fabs(x.x) + __builtin_fabs(x.y); idx = i; }
            for (int off = 32; off > 0; off >>= 1) {
                const double ov = __shfl_xor(v, off, 64);
                const int oi = __shfl_xor(idx, off, 64);
                if (ov > v || (ov == v && oi < idx)) { v = ov; idx = oi; }
            }
            if (threadIdx.x == 0) s_piv = idx;
        }
        __syncthreads();
        const int p = s_piv;
        if (p != k) {
            for (int j = threadIdx.x; j < n; j += blockDim.x) {
                double2 t0 = lM[k * n + j]; lM[k * n + j] = lM[p * n + j]; lM[p * n + j] = t0;
                double2 t1 = lB[k * n + j]; lB[k * n + j] = lB[p * n + j]; lB[p * n + j] = t1;
            }
            __syncthreads();
        }
        const double2 pv = lM[k * n + k];
        const double den = pv.x * pv.x + pv.y * pv.y;
        const double2 ip = make_double2(pv.x / den, -pv.y / den);
        __syncthreads();
        for (int j = threadIdx.x; j < 2 * n; j += blockDim.x) {
            double2* row = j < n ? &lM[k * n + j] : &lB[k * n + (j - n)];
            *row = cmul(*row, ip);
        }
        for (int i = threadIdx.x; i < n; i += blockDim.x) s_f[i] = lM[i * n + k];
        __syncthreads();
        for (int e = threadIdx.x; e < n * 2 * n; e += blockDim.x) {
            const int i = e / (2 * n), j = e % (2 * n);
            if (i == k) continue;
            const double2 f = s_f[i];
            if (j < n) {
                if (j < k) continue;
                const double2 u = cmul(f, lM[k * n + j]);
                lM[i * n + j] = make_double2(lM[i * n + j].x - u.x, lM[i * n + j].y - u.y);
            } else {
                const int c = j - n;
                const double2 u = cmul(f, lB[k * n + c]);
                lB[i * n + c] = make_double2(lB[i * n + c].x - u.x, lB[i * n + c].y - u.y);
            }
        }
        __syncthreads();
    }
    for (int e = threadIdx.x; e < nn; e += blockDim.x) B[e] = lB[e];
    __syncthreads();
}

__global__ __launch_bounds__(kThreads) void k_adj_pade(Dev d, AdjDev a, int c)
{
    __shared__ double2 lA[kNNa], lB[kNNa];
    __shared__ int s_s;
    const int b = blockIdx.x;
    const int m = d.m, n = m + 2, nn = n * n, Bp = d.Bp;
    if (b >= d.B) return;
    double2* W = a.pade_scratch + (size_t)b * a.pade_stride;   // per-trajectory matrices
    double2* As = W;
    double2* A2 = As + kNNa;
    double2* A4 = A2 + kNNa;
    double2* A6 = A4 + kNNa;
    double2* T1 = A6 + kNNa;
    double2* Num = T1 + kNNa;
    double2* Den = Num + kNNa;
    double2* X = Den + kNNa;
    double2* Eb = X + kNNa;
    double2* Tmp = Eb + kNNa;
    double2* Db = Tmp + kNNa;
    double2* A2b = Db + kNNa;
    double2* A4b = A2b + kNNa;
    double2* A6b = A4b + kNNa;
    double2* Asb = A6b + kNNa;
    double2* Es = Asb + kNNa;                                   // E_0 .. E_s
    const double* al = a.alpha + (size_t)c * m * Bp;
    const double* be = a.beta + (size_t)c * m * Bp;
    const double cim = -d.theta;
    auto tb = [&](int i, int j) -> double {
        if (i < m && j < m) {
            if (i == j) return al[(size_t)i * Bp + b];
            if (j == i + 1) return be[(size_t)i * Bp + b];
            if (i == j + 1) return be[(size_t)j * Bp + b];
            return 0.0;
        }
        if (i == 0 && j == m) return 1.0;
        if (i == m && j == m + 1) return 1.0;
        return 0.0;
    };
    if (threadIdx.x == 0) {
        double nrm = 0.0;
        for (int i = 0; i < n; ++i) {
            double rs = 0.0;
            for (int j = 0; j < n; ++j) rs += __builtin_fabs(cim * tb(i, j));
            if (rs > nrm) nrm = rs;
        }
        int s = 0;
        while (nrm > 0.5) { nrm = 0.5 * nrm; s = s + 1; }
        s_s = s;
    }
    __syncthreads();
    const int s = s_s;
    if (s > a.max_s) { if (threadIdx.x == 0) atomicOr(d.status, (int)kStatusScaling); return; }
    const double scl = __builtin_ldexp(1.0, -s);
    double cc[7];
    cc[0] = 1;
    for (int k = 1; k <= 6; ++k) cc[k] = cc[k - 1] * (6.0 + 1. - k) / ((2 * 6.0 + 1. - k) * k);
    for (int e = threadIdx.x; e < nn; e += blockDim.x) {
        const int i = e / n, j = e % n;
        As[e] = make_double2(0.0, (cim * tb(i, j)) * scl);
    }
    __syncthreads();
    gmm(n, As, 0, As, 0, A2, 0, lA, lB);
    gmm(n, A2, 0, A2, 0, A4, 0, lA, lB);
    gmm(n, A2, 0, A4, 0, A6, 0, lA, lB);
    for (int e = threadIdx.x; e < nn; e += blockDim.x) {
        const int i = e / n, j = e % n;
        T1[e] = make_double2(cc[3] * A2[e].x + cc[5] * A4[e].x + (i == j ? cc[1] : 0.0), cc[3] * A2[e].y + cc[5] * A4[e].y);
    }
    __syncthreads();
    gmm(n, As, 0, T1, 0, Num, 0, lA, lB);
    for (int e = threadIdx.x; e < nn; e += blockDim.x) {
        const int i = e / n, j = e % n;
        Den[e] = make_double2(cc[2] * A2[e].x + cc[4] * A4[e].x + cc[6] * A6[e].x - Num[e].x + (i == j ? cc[0] : 0.0),
                              cc[2] * A2[e].y + cc[4] * A4[e].y + cc[6] * A6[e].y - Num[e].y);
        X[e] = Num[e];
    }
    __syncthreads();
    gsolve(n, Den, 0, X, lA, lB);
    for (int e = threadIdx.x; e < nn; e += blockDim.x) {
        const int i = e / n, j = e % n;
        Es[e] = make_double2(2.0 * X[e].x + (i == j ? 1.0 : 0.0), 2.0 * X[e].y);
    }
    __syncthreads();
    for (int k = 0; k < s; ++k) gmm(n, Es + (size_t)k * kNNa, 0, Es + (size_t)k * kNNa, 0, Es + (size_t)(k + 1) * kNNa, 0, lA, lB);
    const double2* E = Es + (size_t)s * kNNa;
    // y_k and the adjoint of E: only column 0, rows < m
    for (int e = threadIdx.x; e < nn; e += blockDim.x) {
        const int i = e / n, j = e % n;
        double2 v = make_double2(0.0, 0.0);
        if (j == 0 && i < m) v = make_double2(d.red[(size_t)(2 * i) * Bp + b], d.red[(size_t)(2 * i + 1) * Bp + b]);
        Eb[e] = v;
        A2b[e] = A4b[e] = A6b[e] = Asb[e] = Db[e] = make_double2(0.0, 0.0);
    }
    for (int k = threadIdx.x; k < m; k += blockDim.x) a.y[(size_t)k * Bp + b] = E[k * n];
    __syncthreads();
    // squaring reverse: Eb_k = Eb E_k^H + E_k^H Eb
    for (int k = s - 1; k >= 0; --k) {
        const double2* Ek = Es + (size_t)k * kNNa;
        gmm(n, Eb, 0, Ek, 1, Tmp, 0, lA, lB);
        gmm(n, Ek, 1, Eb, 0, Tmp, 1, lA, lB);
        for (int e = threadIdx.x; e < nn; e += blockDim.x) Eb[e] = Tmp[e];
        __syncthreads();
    }
    // Xbar = 2 Ebar; Y = Den^-H Xbar (into Tmp); Numbar = Y; Denbar = -Y X^H
    for (int e = threadIdx.x; e < nn; e += blockDim.x) Tmp[e] = make_double2(2.0 * Eb[e].x, 2.0 * Eb[e].y);
    __syncthreads();
    gsolve(n, Den, 1, Tmp, lA, lB);
    gmm(n, Tmp, 0, X, 1, Db, 0, lA, lB);
    double2* Nb = Tmp;
    for (int e = threadIdx.x; e < nn; e += blockDim.x) {
        const double2 db = make_double2(-Db[e].x, -Db[e].y);
        A2b[e] = make_double2(cc[2] * db.x, cc[2] * db.y);
        A4b[e] = make_double2(cc[4] * db.x, cc[4] * db.y);
        A6b[e] = make_double2(cc[6] * db.x, cc[6] * db.y);
        Nb[e] = make_double2(Nb[e].x - db.x, Nb[e].y - db.y);
    }
    __syncthreads();
    // Num = As T1
    gmm(n, Nb, 0, T1, 1, Asb, 1, lA, lB);
    gmm(n, As, 1, Nb, 0, Db, 0, lA, lB);              // T1bar (Db reused)
    for (int e = threadIdx.x; e < nn; e += blockDim.x) {
        A2b[e] = make_double2(A2b[e].x + cc[3] * Db[e].x, A2b[e].y + cc[3] * Db[e].y);
        A4b[e] = make_double2(A4b[e].x + cc[5] * Db[e].x, A4b[e].y + cc[5] * Db[e].y);
    }
    __syncthreads();
    gmm(n, A6b, 0, A4, 1, A2b, 1, lA, lB);            // A6 = A2 A4
    gmm(n, A2, 1, A6b, 0, A4b, 1, lA, lB);
    gmm(n, A4b, 0, A2, 1, A2b, 1, lA, lB);            // A4 = A2 A2
    gmm(n, A2, 1, A4b, 0, A2b, 1, lA, lB);
    gmm(n, A2b, 0, As, 1, Asb, 1, lA, lB);            // A2 = As As
    gmm(n, As, 1, A2b, 0, Asb, 1, lA, lB);
    // Abar = scl Asbar; Tbar_bar = cim Im(Abar); alpha/beta adjoints
    for (int k = threadIdx.x; k < m; k += blockDim.x) {
        a.abar[(size_t)k * Bp + b] = cim * scl * Asb[k * n + k].y;
        a.bbar[(size_t)k * Bp + b] = (k + 1 < m) ? cim * scl * (Asb[k * n + k + 1].y + Asb[(k + 1) * n + k].y) : 0.0;
    }
}

// ---------------------------------------------------------------------------
// three-term reverse, iteration j, part 1: wbar' and partials r2, r3
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kThreads) void k_adj_wbar(Dev d, AdjDev a, int c, int j)
{
    ATile t = adj_tile(d, d.rt_e, d.ntile_e);
    if (!t.ok) return;
    const int b = t.b, Bp = d.Bp, m = d.m;
    const size_t nB = (size_t)d.n * Bp;
    const double2* V = a.basis(c);
    const double* sc = a.sc + (size_t)c * (m + 1) * Bp;
    const double bj = a.beta[((size_t)c * m + j) * Bp + b];
    const double sj1 = sc[(size_t)(j + 1) * Bp + b], sj = sc[(size_t)j * Bp + b];
    const double sjm = j >= 1 ? sc[(size_t)(j - 1) * Bp + b] : 0.0;
    // betabar_j: from the Pade plus -r3 of iteration j+1; r1 = Re<qbar_{j+1}|q_{j+1}>
    double coef_q = 0.0, coef_qb = 0.0;
    if (bj > 0.0 && j + 1 < m) {
        // r1 = Re<qbar_{j+1}|q_{j+1}>: the last reduction (after the SpMM
        // reverse of iteration j+1), still in red[0..Bp)
        const double r1 = d.red[b];
        const double btot = a.bbar[(size_t)j * Bp + b] - a.r3[(size_t)(j + 1) * Bp + b] - r1 / bj;
        coef_qb = 1.0 / bj;
        coef_q = btot * sj1;
    }
    const double2* qb1 = a.qbar_slot(j + 1);
    double acc[2] = {0.0, 0.0};
    for (int i = t.rfirst; i < t.r1; i += 4) {
        const size_t e = (size_t)i * Bp + b;
        double2 w = make_double2(0.0, 0.0);
        if (coef_qb != 0.0) {
            const double2 qb = qb1[e];
            const double2 v1 = V[(size_t)(j + 1) * a.vs + e];
            w = make_double2(coef_qb * qb.x + coef_q * v1.x, coef_qb * qb.y + coef_q * v1.y);
        }
        a.wb[e] = w;
        const double2 vj = V[(size_t)j * a.vs + e];
        acc[0] += sj * (vj.x * w.x + vj.y * w.y);
        if (j >= 1) {
            const double2 vm = V[(size_t)(j - 1) * a.vs + e];
            acc[1] += sjm * (vm.x * w.x + vm.y * w.y);
        }
    }
    adj_store_partials<2>(acc, 2, d.part, d.ntile_e, t.rtile, Bp, b);
}

// per-trajectory scalars after the wbar reduction: r2, r3 (stored per j)
__global__ void k_adj_store_r23(Dev d, AdjDev a, int j)
{
    for (int b = threadIdx.x + blockIdx.x * blockDim.x; b < d.Bp; b += blockDim.x * gridDim.x) {
        a.r2[b] = d.red[b];
        a.r3[(size_t)j * d.Bp + b] = d.red[(size_t)d.Bp + b];
    }
}

// part 2 (gather kernel, one wave per row x 64 trajectories):
//   alphabar = abar_j - r2;  wbar = wbar' + alphabar q_j (also at the gathered rows)
//   qbar_j += -alpha_j wbar' + alphabar (beta_j q_{j+1} + alpha_j q_j + beta_{j-1} q_{j-1}) + H wbar
//   qbar_{j-1} = conj(y_{j-1}) psibar' - beta_{j-1} wbar'
//   gradE += Re(conj(wbar) q_j),  gpart += Re<J wbar|q_j>,  partial r1' = Re<qbar_j|q_j>
__global__ __launch_bounds__(kThreads) void k_adj_spmm(Dev d, AdjDev a, int c, int j)
{
    const int ntt = d.Bp / 64;
    const int logical = xcd_tile(blockIdx.x, gridDim.x);
    if (logical >= d.ntile_s * ntt) return;
    const int rtile = logical / ntt;
    const int lane = threadIdx.x & 63;
    const int wave = uniform(threadIdx.x >> 6);
    const int b = (logical % ntt) * 64 + lane;
    const int Bp = d.Bp, m = d.m;
    const size_t nB = (size_t)d.n * Bp;
    const double2* V = a.basis(c);
    const double* sc = a.sc + (size_t)c * (m + 1) * Bp;
    const double* al = a.alpha + (size_t)c * m * Bp;
    const double* be = a.beta + (size_t)c * m * Bp;
    const double* diag = a.diag_step(c);
    const double sj = sc[(size_t)j * Bp + b];
    const double sj1 = sc[(size_t)(j + 1) * Bp + b];
    const double sjm = j >= 1 ? sc[(size_t)(j - 1) * Bp + b] : 0.0;
    const double aj = al[(size_t)j * Bp + b];
    const double bj = be[(size_t)j * Bp + b];
    const double bjm = j >= 1 ? be[(size_t)(j - 1) * Bp + b] : 0.0;
    const double abart = a.abar[(size_t)j * Bp + b] - a.r2[b];
    const double2 yj = a.y[(size_t)j * Bp + b];
    const double2 yjm = j >= 1 ? a.y[(size_t)(j - 1) * Bp + b] : make_double2(0.0, 0.0);
    const double2* vj = V + (size_t)j * a.vs;
    double2* qbj = a.qbar_slot(j);
    double2* qbm = j >= 1 ? a.qbar_slot(j - 1) : nullptr;
    const double cq = abart * sj;                 // wbar = wbar' + cq v_j
    const double g = d.g;
    double accg = 0.0, accr = 0.0;
    const int r0 = rtile * d.rt_s, r1 = min(d.n, r0 + d.rt_s);
    for (int i = r0 + wave; i < r1; i += 4) {
        const int p0 = d.rowptr[i], p1 = d.rowptr[i + 1];
        double sr = 0.0, si = 0.0;
        for (int p = p0; p < p1; ++p) {
            const int cidx = d.col[p];
            const double v = d.val64[p];
            const size_t ec = (size_t)cidx * Bp + b;
            const double2 w1 = a.wb[ec], x1 = vj[ec];
            sr += v * (w1.x + cq * x1.x);
            si += v * (w1.y + cq * x1.y);
        }
        const size_t e = (size_t)i * Bp + b;
        const double2 wp = a.wb[e];
        const double2 xj = vj[e];
        const double2 q = make_double2(sj * xj.x, sj * xj.y);
        const double2 wbar = make_double2(wp.x + abart * q.x, wp.y + abart * q.y);
        const double di = diag[e];
        const double2 Hw = make_double2(di * wbar.x + g * sr, di * wbar.y + g * si);
        // H q_j by the three-term relation
        const double2 x1 = V[(size_t)(j + 1) * a.vs + e];
        double2 hq = make_double2(bj * sj1 * x1.x + aj * q.x, bj * sj1 * x1.y + aj * q.y);
        if (j >= 1) {
            const double2 xm = V[(size_t)(j - 1) * a.vs + e];
            hq.x += bjm * sjm * xm.x;
            hq.y += bjm * sjm * xm.y;
        }
        double2 qb = qbj[e];
        qb.x += -aj * wp.x + abart * hq.x + Hw.x;
        qb.y += -aj * wp.y + abart * hq.y + Hw.y;
        qbj[e] = qb;
        if (j >= 1) {
            const double2 pb = a.pbar1[e];
            // conj(y_{j-1}) psibar'
            qbm[e] = make_double2(yjm.x * pb.x + yjm.y * pb.y - bjm * wp.x, yjm.x * pb.y - yjm.y * pb.x - bjm * wp.y);
        }
        a.grad_e[e] += wbar.x * q.x + wbar.y * q.y;
        accg += sr * q.x + si * q.y;
        accr += qb.x * q.x + qb.y * q.y;
    }
    (void)yj;
    // gbar accumulates in place over the whole sweep (each tile owns its slots)
    __shared__ double lds[4][64];
    __syncthreads();
    lds[wave][lane] = accg;
    __syncthreads();
    if (wave == 0) a.gpart[(size_t)rtile * Bp + b] += (lds[0][lane] + lds[1][lane]) + (lds[2][lane] + lds[3][lane]);
    double acc[1] = {accr};
    adj_store_partials<1>(acc, 1, d.part, d.ntile_s, rtile, Bp, b);
}

// Row-blocked, fused form of k_adj_wbar + k_adj_spmm (g != 0).  A wavefront
// walks the union columns of kRB rows and gathers qbar_{j+1} only:
//   wbar' = qbar_{j+1}/beta_j + btot q_{j+1}
//     (btot = betabar_j - Re<q_j|wbar'_{j+1}> - Re<qbar_{j+1}|q_{j+1}>/beta_j),
//   J wbar' = J qbar_{j+1}/beta_j + btot J q_{j+1},
// with J q_k = (H q_k - d q_k)/g from the three-term relation
// H q_k = beta_k q_{k+1} + alpha_k q_k + beta_{k-1} q_{k-1}, the relation the
// reverse already uses for alphabar (H q_j).  The global sums this needs come
// from the previous iteration's reduction, red[0..4):
//   Re<qbar_{j+1}|q_{j+1}>, Re<q_j|qbar_{j+1}>, Re<q_j|q_{j+1}>, Re<q_j|wbar'_{j+1}>
// (so r2 = Re<q_j|wbar'> = Re<q_j|qbar_{j+1}>/beta_j + btot Re<q_j|q_{j+1}>),
// and this kernel writes the same four partial sums for iteration j-1.  Rows
// at or beyond a trajectory's n_active take no update, as the forward SpMM
// writes 0 there.
template <int kRB>
__global__ __launch_bounds__(kThreads) void k_adj_spmm_rb(Dev d, AdjDev a, int c, int j)
{
    const int ntt = d.Bp / 64;
    const int logical = xcd_tile(blockIdx.x, gridDim.x);
    if (logical >= d.ntile_s * ntt) return;
    const int rtile = logical / ntt;
    const int lane = threadIdx.x & 63;
    const int wave = uniform(threadIdx.x >> 6);
    const int slice = logical % ntt;
    const int b = slice * 64 + lane;
    const int Bp = d.Bp, m = d.m;
    const size_t nB = (size_t)d.n * Bp;
    const double2* V = a.basis(c);
    const double* sc = a.sc + (size_t)c * (m + 1) * Bp;
    const double* al = a.alpha + (size_t)c * m * Bp;
    const double* be = a.beta + (size_t)c * m * Bp;
    const double* diag = a.diag_step(c);
    auto S = [&](int k) { return (k >= 0 && k <= m) ? sc[(size_t)k * Bp + b] : 0.0; };
    auto A = [&](int k) { return (k >= 0 && k < m) ? al[(size_t)k * Bp + b] : 0.0; };
    auto Bt = [&](int k) { return (k >= 0 && k < m) ? be[(size_t)k * Bp + b] : 0.0; };
    const double sj = S(j), sj1 = S(j + 1), sj2 = S(j + 2), sjm = S(j - 1);
    const double aj = A(j), aj1 = A(j + 1);
    const double bj = Bt(j), bj1 = Bt(j + 1), bjm = Bt(j - 1);
    // wbar' = cqb qbar_{j+1} + btot q_{j+1}
    const bool top = !(bj > 0.0 && j + 1 < m);
    double cqb = 0.0, btot = 0.0, r2 = 0.0;
    if (!top) {
        const double R0 = d.red[b], R1 = d.red[(size_t)Bp + b], R2 = d.red[(size_t)2 * Bp + b];
        const double R3 = d.red[(size_t)3 * Bp + b];
        cqb = 1.0 / bj;
        btot = a.bbar[(size_t)j * Bp + b] - R3 - R0 / bj;
        r2 = cqb * R1 + btot * R2;
    }
    const double abart = a.abar[(size_t)j * Bp + b] - r2;
    const double2 yjm = j >= 1 ? a.y[(size_t)(j - 1) * Bp + b] : make_double2(0.0, 0.0);
    double2* qbj = a.qbar_slot(j);
    double2* qbm = j >= 1 ? a.qbar_slot(j - 1) : nullptr;
    const double2* qb1 = a.qbar_slot(j + 1);
    const double g = d.g, ginv = 1.0 / d.g;
    const int nact = d.nact[b];
    const double2* __restrict__ xs = qb1 + (size_t)slice * 64;
    const double* __restrict__ rv = d.rb_val64;
    const int32_t* __restrict__ rc = d.rb_col;
    const int nblk = (d.n + kRB - 1) / kRB;
    const int rb = rtile * 4 + wave;
    double acc[5] = {0.0, 0.0, 0.0, 0.0, 0.0};   // the four sums for j-1, gbar
    if (rb < nblk) {
        double sr[kRB], si[kRB];
#pragma unroll
        for (int r = 0; r < kRB; ++r) { sr[r] = 0.0; si[r] = 0.0; }
        if (!uniform((int)top)) {
            const int u0 = uniform(d.rb_ptr[rb]), u1 = uniform(d.rb_ptr[rb + 1]);
            int p = u0;
            for (; p + 8 <= u1; p += 8) {
                double2 xw[8];
#pragma unroll
                for (int u = 0; u < 8; ++u) xw[u] = xs[(size_t)uniform(rc[p + u]) * Bp + lane];
#pragma unroll
                for (int u = 0; u < 8; ++u) {
#pragma unroll
                    for (int r = 0; r < kRB; ++r) {
                        const double v = rv[(size_t)(p + u) * kRB + r];
                        sr[r] = __builtin_fma(v, xw[u].x, sr[r]);
                        si[r] = __builtin_fma(v, xw[u].y, si[r]);
                    }
                }
            }
            for (; p < u1; ++p) {
                const double2 xw = xs[(size_t)uniform(rc[p]) * Bp + lane];
#pragma unroll
                for (int r = 0; r < kRB; ++r) {
                    const double v = rv[(size_t)p * kRB + r];
                    sr[r] = __builtin_fma(v, xw.x, sr[r]);
                    si[r] = __builtin_fma(v, xw.y, si[r]);
                }
            }
        }
#pragma unroll
        for (int r = 0; r < kRB; ++r) {
            const int i = rb * kRB + r;
            if (i >= d.n || i >= nact) continue;
            const size_t e = (size_t)i * Bp + b;
            const double di = diag[e];
            const double2 xj = V[(size_t)j * a.vs + e];
            const double2 q = make_double2(sj * xj.x, sj * xj.y);
            const double2 x1 = V[(size_t)(j + 1) * a.vs + e];
            const double2 q1 = make_double2(sj1 * x1.x, sj1 * x1.y);
            double2 qm = make_double2(0.0, 0.0);
            if (j >= 1) { const double2 t = V[(size_t)(j - 1) * a.vs + e]; qm = make_double2(sjm * t.x, sjm * t.y); }
            // H q_j by the three-term relation
            const double2 hq = make_double2(bj * q1.x + aj * q.x + bjm * qm.x, bj * q1.y + aj * q.y + bjm * qm.y);
            double2 wp = make_double2(0.0, 0.0), Jw = make_double2(0.0, 0.0);
            if (!top) {
                const double2 qbn = qb1[e];
                wp = make_double2(cqb * qbn.x + btot * q1.x, cqb * qbn.y + btot * q1.y);
                // H q_{j+1} by the three-term relation (slot j+2 <= m holds q_{j+2})
                const double2 x2 = V[(size_t)(j + 2) * a.vs + e];
                const double2 hq1 = make_double2(bj1 * sj2 * x2.x + aj1 * q1.x + bj * q.x,
                                                 bj1 * sj2 * x2.y + aj1 * q1.y + bj * q.y);
                Jw = make_double2(cqb * sr[r] + btot * (hq1.x - di * q1.x) * ginv,
                                  cqb * si[r] + btot * (hq1.y - di * q1.y) * ginv);
            }
            const double2 wbar = make_double2(wp.x + abart * q.x, wp.y + abart * q.y);
            // J wbar = J wbar' + alphabar J q_j
            Jw.x += abart * (hq.x - di * q.x) * ginv;
            Jw.y += abart * (hq.y - di * q.y) * ginv;
            const double2 Hw = make_double2(di * wbar.x + g * Jw.x, di * wbar.y + g * Jw.y);
            double2 qb = qbj[e];
            qb.x += -aj * wp.x + abart * hq.x + Hw.x;
            qb.y += -aj * wp.y + abart * hq.y + Hw.y;
            qbj[e] = qb;
            if (j >= 1) {
                const double2 pb = a.pbar1[e];
                qbm[e] = make_double2(yjm.x * pb.x + yjm.y * pb.y - bjm * wp.x, yjm.x * pb.y - yjm.y * pb.x - bjm * wp.y);
            }
            a.grad_e[e] += wbar.x * q.x + wbar.y * q.y;
            acc[0] += qb.x * q.x + qb.y * q.y;              // Re<qbar_j|q_j>
            acc[1] += qm.x * qb.x + qm.y * qb.y;            // Re<q_{j-1}|qbar_j>
            acc[2] += qm.x * q.x + qm.y * q.y;              // Re<q_{j-1}|q_j>
            acc[3] += qm.x * wp.x + qm.y * wp.y;            // Re<q_{j-1}|wbar'_j>
            acc[4] += Jw.x * q.x + Jw.y * q.y;              // Re<J wbar|q_j>
        }
    }
    __shared__ double lds[4][64];
    __syncthreads();
    lds[wave][lane] = acc[4];
    __syncthreads();
    if (wave == 0) a.gpart[(size_t)rtile * Bp + b] += (lds[0][lane] + lds[1][lane]) + (lds[2][lane] + lds[3][lane]);
    adj_store_partials<4>(*reinterpret_cast<const double(*)[4]>(acc), 4, d.part, d.ntile_s, rtile, Bp, b);
}

// qbar_{m-1} = conj(y_{m-1}) psibar' (start of the three-term reverse)
__global__ void k_adj_init_top(Dev d, AdjDev a)
{
    const int m = d.m, Bp = d.Bp;
    const size_t nB = (size_t)d.n * Bp;
    double2* qb = a.qbar_slot(m - 1);
    for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < nB; e += (size_t)gridDim.x * blockDim.x) {
        const int b = (int)(e % Bp);
        const double2 y = a.y[(size_t)(m - 1) * Bp + b], pb = a.pbar1[e];
        qb[e] = make_double2(y.x * pb.x + y.y * pb.y, y.x * pb.y - y.y * pb.x);
    }
}

// pbar += bra conj? -- C(n) = <bra|psi(n)>: psibar += bra * Cbar(n)
__global__ void k_adj_add_corr(Dev d, AdjDev a, int64_t step, double2* dst)
{
    const int Bp = d.Bp;
    const size_t nB = (size_t)d.n * Bp;
    for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < nB; e += (size_t)gridDim.x * blockDim.x) {
        const int i = (int)(e / Bp), b = (int)(e % Bp);
        const double2 br = d.bra[(d.bra_idx ? (size_t)d.bra_idx[b] * d.n : 0) + i], cb = a.cbar[(size_t)step * Bp + b];
        double2 p = dst[e];
        p.x += br.x * cb.x - br.y * cb.y;
        p.y += br.x * cb.y + br.y * cb.x;
        dst[e] = p;
    }
}

// ---------------------------------------------------------------------------
hipError_t adj_renorm(const Dev& d, const AdjDev& a, int c, hipStream_t s)
{
    const int ntt = d.Bp / 64;
    dim3 grid(pad8a(d.ntile_e * ntt)), block(kThreads);
    hipLaunchKernelGGL(k_adj_renorm_dot, grid, block, 0, s, d, a, c);
    hipError_t e = launch_reduce(d, 1, d.ntile_e, s);
    if (e != hipSuccess) return e;
    const dim3 g16(pad8a(d.ntile_e * (d.Bp / 16)));
    if (d.m <= 8) hipLaunchKernelGGL((k_adj_renorm_apply<2>), g16, block, 0, s, d, a, c);
    else if (d.m <= 16) hipLaunchKernelGGL((k_adj_renorm_apply<4>), g16, block, 0, s, d, a, c);
    else if (d.m <= 24) hipLaunchKernelGGL((k_adj_renorm_apply<6>), g16, block, 0, s, d, a, c);
    else hipLaunchKernelGGL((k_adj_renorm_apply<8>), g16, block, 0, s, d, a, c);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    return launch_reduce(d, 2 * d.m, d.ntile_e, s);
}

hipError_t adj_pade(const Dev& d, const AdjDev& a, int c, hipStream_t s)
{
    hipLaunchKernelGGL(k_adj_pade, dim3(d.Bp), dim3(kThreads), 0, s, d, a, c);
    return hipGetLastError();
}

hipError_t adj_lanczos(const Dev& d, const AdjDev& a, int c, hipStream_t s)
{
    const int ntt = d.Bp / 64;
    dim3 ge(pad8a(d.ntile_e * ntt)), gs(pad8a(d.ntile_s * ntt)), block(kThreads);
    const size_t nB = (size_t)d.n * d.Bp;
    int gi = (int)std::min<size_t>((nB + kThreads - 1) / kThreads, 65535);
    hipLaunchKernelGGL(k_adj_init_top, dim3(gi), block, 0, s, d, a);
    hipError_t e;
    if (d.g != 0.0) {
        // fused: one gather kernel and one reduction per iteration
        for (int j = d.m - 1; j >= 0; --j) {
            switch (d.rb) {
            case 2: hipLaunchKernelGGL((k_adj_spmm_rb<2>), gs, block, 0, s, d, a, c, j); break;
            case 8: hipLaunchKernelGGL((k_adj_spmm_rb<8>), gs, block, 0, s, d, a, c, j); break;
            default: hipLaunchKernelGGL((k_adj_spmm_rb<4>), gs, block, 0, s, d, a, c, j); break;
            }
            if ((e = launch_reduce_wide(d, 4, d.ntile_s, 0, s)) != hipSuccess) return e;
        }
        return hipGetLastError();
    }
    for (int j = d.m - 1; j >= 0; --j) {
        hipLaunchKernelGGL(k_adj_wbar, ge, block, 0, s, d, a, c, j);
        if ((e = launch_reduce(d, 2, d.ntile_e, s)) != hipSuccess) return e;
        hipLaunchKernelGGL(k_adj_store_r23, dim3(1), block, 0, s, d, a, j);
        hipLaunchKernelGGL(k_adj_spmm, gs, block, 0, s, d, a, c, j);
        if ((e = launch_reduce(d, 1, d.ntile_s, s)) != hipSuccess) return e;
    }
    return hipGetLastError();
}

hipError_t adj_add_corr(const Dev& d, const AdjDev& a, int64_t step, double2* dst, hipStream_t s)
{
    const size_t nB = (size_t)d.n * d.Bp;
    int gi = (int)std::min<size_t>((nB + kThreads - 1) / kThreads, 65535);
    hipLaunchKernelGGL(k_adj_add_corr, dim3(gi), dim3(kThreads), 0, s, d, a, step, dst);
    return hipGetLastError();
}

}  // namespace dq
