/*
 * adjoint.c -- oracle reverse-mode gradient of a trajectory loss.
 * TEST INFRASTRUCTURE.
 *
 * The reference has no derivatives.  This differentiates the discrete map
 * of krylov.c exactly (not the continuous propagator), so it agrees with
 * central finite differences of orc_run whatever the Krylov error:
 *   noise (parameter independent, replayed from the Philox stream),
 *   per cycle: Lanczos with full CGS projection  w = H q_j,
 *     h_k = q_k^H w (k <= j), alpha_j = Re h_j, w' = w - sum h_k q_k,
 *     beta_j = |w'|, q_{j+1} = w'/beta_j   (lanczosClass.cu:395-531),
 *   Pade(6,6) with scaling and squaring of A = -i theta Tbar
 *     (padeClass.cu:204-706; s is held fixed, it is piecewise constant),
 *   psi' = sum_k y_k q_k (propStep.cu:31-44), psi'' = psi'/|psi'|
 *     (systemClass.cu:1769-1829).
 * The oracle's GRAM projections equal the CGS ones in exact arithmetic, so
 * this is the derivative of both.  Loss:
 *   L = sum_g w_pop[g] P_g(T) + w_spec sum_{n=0..T} u(n) |C(n) - C_target(n)|^2,
 *   P_g = sum_{i in g} |psi_i|^2,  C(n) = <bra|psi(n)>.
 * Parameters: the site energies (additive on the diagonal at every step)
 * and the coupling scale g.  States are checkpointed every `ckpt` steps;
 * each segment is recomputed from its checkpoint with the noise replayed.
 */
#include "dq_oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>

typedef struct { double re, im; } cx;

/* ---------------- small complex dense helpers (column-major) ------------ */
static void cmm(int n, const double *A, const double *B, double *C)
{
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i) {
            double sr = 0.0, si = 0.0;
            for (int k = 0; k < n; ++k) {
                const double *a = A + 2 * (i + k * n), *b = B + 2 * (k + j * n);
                sr += a[0] * b[0] - a[1] * b[1];
                si += a[0] * b[1] + a[1] * b[0];
            }
            C[2 * (i + j * n)] = sr;
            C[2 * (i + j * n) + 1] = si;
        }
}

/* C += A^H B */
static void cmm_ha_acc(int n, const double *A, const double *B, double *C)
{
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i) {
            double sr = 0.0, si = 0.0;
            for (int k = 0; k < n; ++k) {
                const double *a = A + 2 * (k + i * n), *b = B + 2 * (k + j * n);
                sr += a[0] * b[0] + a[1] * b[1];
                si += a[0] * b[1] - a[1] * b[0];
            }
            C[2 * (i + j * n)] += sr;
            C[2 * (i + j * n) + 1] += si;
        }
}

/* C += A B^H */
static void cmm_bh_acc(int n, const double *A, const double *B, double *C)
{
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i) {
            double sr = 0.0, si = 0.0;
            for (int k = 0; k < n; ++k) {
                const double *a = A + 2 * (i + k * n), *b = B + 2 * (j + k * n);
                sr += a[0] * b[0] + a[1] * b[1];
                si += a[1] * b[0] - a[0] * b[1];
            }
            C[2 * (i + j * n)] += sr;
            C[2 * (i + j * n) + 1] += si;
        }
}

/* solve M X = B in place (Gaussian elimination, partial pivoting) */
static int csolve(int n, double *M, double *B)
{
    for (int k = 0; k < n; ++k) {
        int p = k;
        double best = -1.0;
        for (int i = k; i < n; ++i) {
            double v = fabs(M[2 * (i + k * n)]) + fabs(M[2 * (i + k * n) + 1]);
            if (v > best) { best = v; p = i; }
        }
        if (!(best > 0.0)) return -1;
        if (p != k)
            for (int j = 0; j < n; ++j) {
                for (int t = 0; t < 2; ++t) {
                    double x = M[2 * (k + j * n) + t]; M[2 * (k + j * n) + t] = M[2 * (p + j * n) + t]; M[2 * (p + j * n) + t] = x;
                    x = B[2 * (k + j * n) + t]; B[2 * (k + j * n) + t] = B[2 * (p + j * n) + t]; B[2 * (p + j * n) + t] = x;
                }
            }
        double pr = M[2 * (k + k * n)], pi = M[2 * (k + k * n) + 1];
        double den = pr * pr + pi * pi, ir = pr / den, ii = -pi / den;
        for (int i = k + 1; i < n; ++i) {
            double *a = M + 2 * (i + k * n);
            double lr = a[0] * ir - a[1] * ii, li = a[0] * ii + a[1] * ir;
            for (int j = k + 1; j < n; ++j) {
                double *d = M + 2 * (i + j * n);
                const double *u = M + 2 * (k + j * n);
                d[0] -= lr * u[0] - li * u[1];
                d[1] -= lr * u[1] + li * u[0];
            }
            for (int j = 0; j < n; ++j) {
                double *d = B + 2 * (i + j * n);
                const double *u = B + 2 * (k + j * n);
                d[0] -= lr * u[0] - li * u[1];
                d[1] -= lr * u[1] + li * u[0];
            }
        }
    }
    for (int j = 0; j < n; ++j)
        for (int i = n - 1; i >= 0; --i) {
            double sr = B[2 * (i + j * n)], si = B[2 * (i + j * n) + 1];
            for (int k = i + 1; k < n; ++k) {
                const double *u = M + 2 * (i + k * n), *x = B + 2 * (k + j * n);
                sr -= u[0] * x[0] - u[1] * x[1];
                si -= u[0] * x[1] + u[1] * x[0];
            }
            double pr = M[2 * (i + i * n)], pi = M[2 * (i + i * n) + 1];
            double den = pr * pr + pi * pi;
            B[2 * (i + j * n)] = (sr * pr + si * pi) / den;
            B[2 * (i + j * n) + 1] = (si * pr - sr * pi) / den;
        }
    return 0;
}

/* Pade forward keeping what the reverse sweep needs */
typedef struct {
    int n, s;
    double sc;
    double *As, *A2, *A4, *A6, *T1, *Num, *Den, *X, *Es; /* Es: s+1 matrices E_0..E_s */
} pade_tape;

static int pade_forward(int n, const double *A, pade_tape *t)
{
    const size_t sz = (size_t)2 * n * n;
    double c[7];
    orc_pade_coeffs(6, c);
    double nrm = 0.0;
    for (int i = 0; i < n; ++i) {
        double rs = 0.0;
        for (int j = 0; j < n; ++j) {
            double re = A[2 * (i + j * n)], im = A[2 * (i + j * n) + 1];
            rs += sqrt(re * re + im * im);
        }
        if (rs > nrm) nrm = rs;
    }
    int s = 0;
    while (nrm > 0.5) { nrm = 0.5 * nrm; s = s + 1; }
    if (s >= 50) return -1;
    t->n = n;
    t->s = s;
    t->sc = ldexp(1.0, -s);
    double *buf = (double *)malloc(sizeof(double) * sz * (8 + s + 1));
    t->As = buf; t->A2 = buf + sz; t->A4 = buf + 2 * sz; t->A6 = buf + 3 * sz;
    t->T1 = buf + 4 * sz; t->Num = buf + 5 * sz; t->Den = buf + 6 * sz; t->X = buf + 7 * sz;
    t->Es = buf + 8 * sz;
    for (size_t i = 0; i < sz; ++i) t->As[i] = A[i] * t->sc;
    cmm(n, t->As, t->As, t->A2);
    cmm(n, t->A2, t->A2, t->A4);
    cmm(n, t->A2, t->A4, t->A6);
    for (size_t i = 0; i < sz; ++i) t->T1[i] = c[3] * t->A2[i] + c[5] * t->A4[i];
    for (int i = 0; i < n; ++i) t->T1[2 * (i + i * n)] += c[1];
    cmm(n, t->As, t->T1, t->Num);
    for (size_t i = 0; i < sz; ++i)
        t->Den[i] = c[2] * t->A2[i] + c[4] * t->A4[i] + c[6] * t->A6[i] - t->Num[i];
    for (int i = 0; i < n; ++i) t->Den[2 * (i + i * n)] += c[0];
    double *M = (double *)malloc(sizeof(double) * sz);
    memcpy(M, t->Den, sz * sizeof(double));
    memcpy(t->X, t->Num, sz * sizeof(double));
    if (csolve(n, M, t->X) != 0) { free(M); free(buf); return -2; }
    free(M);
    for (size_t i = 0; i < sz; ++i) t->Es[i] = 2.0 * t->X[i];
    for (int i = 0; i < n; ++i) t->Es[2 * (i + i * n)] += 1.0;
    for (int k = 0; k < s; ++k) cmm(n, t->Es + k * sz, t->Es + k * sz, t->Es + (k + 1) * sz);
    return 0;
}

/* Ebar (n x n) -> Abar */
static void pade_reverse(pade_tape *t, const double *Ebar_in, double *Abar)
{
    const int n = t->n;
    const size_t sz = (size_t)2 * n * n;
    double c[7];
    orc_pade_coeffs(6, c);
    double *w = (double *)calloc(sz * 9, sizeof(double));
    double *Eb = w, *tmp = w + sz, *Xb = w + 2 * sz, *Nb = w + 3 * sz, *Db = w + 4 * sz;
    double *A2b = w + 5 * sz, *A4b = w + 6 * sz, *A6b = w + 7 * sz, *Asb = w + 8 * sz;
    memcpy(Eb, Ebar_in, sz * sizeof(double));
    for (int k = t->s - 1; k >= 0; --k) {
        const double *Ek = t->Es + k * sz;
        memset(tmp, 0, sz * sizeof(double));
        cmm_bh_acc(n, Eb, Ek, tmp);   /* Eb Ek^H */
        cmm_ha_acc(n, Ek, Eb, tmp);   /* Ek^H Eb */
        memcpy(Eb, tmp, sz * sizeof(double));
    }
    for (size_t i = 0; i < sz; ++i) Xb[i] = 2.0 * Eb[i];
    /* X = Den^-1 Num:  Y = Den^-H Xb;  Numbar = Y;  Denbar = -Y X^H */
    double *M = (double *)malloc(sizeof(double) * sz);
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) {
            M[2 * (i + j * n)] = t->Den[2 * (j + i * n)];
            M[2 * (i + j * n) + 1] = -t->Den[2 * (j + i * n) + 1];
        }
    memcpy(Nb, Xb, sz * sizeof(double));
    csolve(n, M, Nb);
    free(M);
    cmm_bh_acc(n, Nb, t->X, Db);
    for (size_t i = 0; i < sz; ++i) Db[i] = -Db[i];
    /* Den = c2 A2 + c4 A4 + c6 A6 - Num + c0 I */
    for (size_t i = 0; i < sz; ++i) {
        A2b[i] += c[2] * Db[i];
        A4b[i] += c[4] * Db[i];
        A6b[i] += c[6] * Db[i];
        Nb[i] -= Db[i];
    }
    /* Num = As T1 */
    cmm_bh_acc(n, Nb, t->T1, Asb);
    memset(tmp, 0, sz * sizeof(double));
    cmm_ha_acc(n, t->As, Nb, tmp);          /* T1bar */
    for (size_t i = 0; i < sz; ++i) { A2b[i] += c[3] * tmp[i]; A4b[i] += c[5] * tmp[i]; }
    /* A6 = A2 A4 */
    cmm_bh_acc(n, A6b, t->A4, A2b);
    cmm_ha_acc(n, t->A2, A6b, A4b);
    /* A4 = A2 A2 */
    cmm_bh_acc(n, A4b, t->A2, A2b);
    cmm_ha_acc(n, t->A2, A4b, A2b);
    /* A2 = As As */
    cmm_bh_acc(n, A2b, t->As, Asb);
    cmm_ha_acc(n, t->As, A2b, Asb);
    for (size_t i = 0; i < sz; ++i) Abar[i] = t->sc * Asb[i];
    free(w);
}

/* --------------------------- one Krylov cycle ---------------------------- */
typedef struct {
    int n, m;
    const int32_t *rowptr, *col;
    const double *val;
    double g;
    double theta;          /* tau / hbar */
    double *Q;             /* (m+2) x 2n */
    double *W;             /* w = H q_j, (m+1) x 2n */
    double *h;             /* (m+1) x (m+1) complex: h[j][k] */
    double *alpha, *beta;  /* m+2 */
    int jb;                /* iterations completed (m, or less at breakdown) */
    double *y, *psi1, nrm; /* combine */
    pade_tape tape;
    double *work;          /* 2n */
} cycle_t;

static void jv(const cycle_t *c, const double *x, double *y)   /* y = J x */
{
    for (int i = 0; i < c->n; ++i) {
        double ar = 0.0, ai = 0.0;
        for (int32_t p = c->rowptr[i]; p < c->rowptr[i + 1]; ++p) {
            ar += c->val[p] * x[2 * c->col[p]];
            ai += c->val[p] * x[2 * c->col[p] + 1];
        }
        y[2 * i] = ar;
        y[2 * i + 1] = ai;
    }
}

static void cycle_alloc(cycle_t *c, int n, int m)
{
    c->n = n; c->m = m;
    c->Q = (double *)calloc((size_t)2 * n * (m + 2), sizeof(double));
    c->W = (double *)calloc((size_t)2 * n * (m + 2), sizeof(double));
    c->h = (double *)calloc((size_t)2 * (m + 2) * (m + 2), sizeof(double));
    c->alpha = (double *)calloc(m + 2, sizeof(double));
    c->beta = (double *)calloc(m + 2, sizeof(double));
    c->y = (double *)calloc((size_t)2 * m, sizeof(double));
    c->psi1 = (double *)calloc((size_t)2 * n, sizeof(double));
    c->work = (double *)calloc((size_t)2 * n, sizeof(double));
    c->tape.As = NULL;
}

static void cycle_free(cycle_t *c)
{
    free(c->Q); free(c->W); free(c->h); free(c->alpha); free(c->beta);
    free(c->y); free(c->psi1); free(c->work);
    if (c->tape.As) free(c->tape.As);
}

#define HJ(c, j, k) ((c)->h + 2 * ((size_t)(j) * ((c)->m + 2) + (k)))

/* forward of one cycle from psi; returns 0 or <0 */
static int cycle_forward(cycle_t *c, const double *diag, const double *psi, double tol)
{
    const int n = c->n, m = c->m;
    const size_t cs = 2 * (size_t)n;
    memset(c->Q, 0, cs * (m + 2) * sizeof(double));
    memcpy(c->Q + cs, psi, cs * sizeof(double));
    memset(c->alpha, 0, (m + 2) * sizeof(double));
    memset(c->beta, 0, (m + 2) * sizeof(double));
    c->jb = m;
    for (int j = 1; j <= m; ++j) {
        const double *q = c->Q + cs * j;
        double *w = c->W + cs * j;
        orc_spmv(n, c->rowptr, c->col, c->val, c->g, diag, q, w);
        double *wp = c->work;
        memcpy(wp, w, cs * sizeof(double));
        for (int k = 1; k <= j; ++k) {
            const double *qk = c->Q + cs * k;
            double sr = 0.0, si = 0.0;
            for (int i = 0; i < n; ++i) {
                sr += qk[2 * i] * w[2 * i] + qk[2 * i + 1] * w[2 * i + 1];
                si += qk[2 * i] * w[2 * i + 1] - qk[2 * i + 1] * w[2 * i];
            }
            HJ(c, j, k)[0] = sr;
            HJ(c, j, k)[1] = si;
        }
        c->alpha[j] = HJ(c, j, j)[0];
        for (int k = 1; k <= j; ++k) {
            const double *qk = c->Q + cs * k;
            double hr = HJ(c, j, k)[0], hi = HJ(c, j, k)[1];
            for (int i = 0; i < n; ++i) {
                wp[2 * i] -= hr * qk[2 * i] - hi * qk[2 * i + 1];
                wp[2 * i + 1] -= hr * qk[2 * i + 1] + hi * qk[2 * i];
            }
        }
        double b = 0.0;
        for (size_t i = 0; i < cs; ++i) b += wp[i] * wp[i];
        b = sqrt(b);
        if (!(b > tol * (fabs(c->alpha[j]) + c->beta[j - 1]))) { c->beta[j] = 0.0; c->jb = j; break; }
        c->beta[j] = b;
        double *qn = c->Q + cs * (j + 1);
        for (size_t i = 0; i < cs; ++i) qn[i] = wp[i] / b;
    }
    /* Pade of A = -i theta Tbar */
    const int nt = m + 2;
    double *tb = (double *)malloc(sizeof(double) * nt * nt);
    double *A = (double *)malloc(sizeof(double) * 2 * nt * nt);
    orc_build_tbar(m, c->alpha, c->beta, tb);
    double cim = -c->theta;
    for (int i = 0; i < nt * nt; ++i) { A[2 * i] = 0.0; A[2 * i + 1] = cim * tb[i]; }
    if (c->tape.As) { free(c->tape.As); c->tape.As = NULL; }
    int rc = pade_forward(nt, A, &c->tape);
    free(tb); free(A);
    if (rc != 0) return -1;
    const double *E = c->tape.Es + (size_t)c->tape.s * 2 * nt * nt;
    for (int k = 0; k < m; ++k) { c->y[2 * k] = E[2 * k]; c->y[2 * k + 1] = E[2 * k + 1]; }
    memset(c->psi1, 0, cs * sizeof(double));
    for (int k = 1; k <= m; ++k) {
        const double *q = c->Q + cs * k;
        double yr = c->y[2 * (k - 1)], yi = c->y[2 * (k - 1) + 1];
        for (int i = 0; i < n; ++i) {
            c->psi1[2 * i] += yr * q[2 * i] - yi * q[2 * i + 1];
            c->psi1[2 * i + 1] += yr * q[2 * i + 1] + yi * q[2 * i];
        }
    }
    double nr = 0.0;
    for (size_t i = 0; i < cs; ++i) nr += c->psi1[i] * c->psi1[i];
    c->nrm = sqrt(nr);
    return 0;
}

/* Reverse of the Lanczos part by the three-term recurrence
 * w' = H q_j - alpha_j q_j - beta_{j-1} q_{j-1}, using the stored
 * (reorthogonalised) basis and H q_j = beta_j q_{j+1} + alpha_j q_j +
 * beta_{j-1} q_{j-1}.  The full-projection map and the three-term map are
 * the same function of psi in exact arithmetic, so their derivatives agree
 * up to the basis' loss of orthogonality; this costs O(m) vector operations
 * per cycle instead of O(m^2).  Qb[k] holds the adjoints of q_k from the
 * combine; ab/bb the adjoints of alpha/beta from the Pade. */
static void lanczos_reverse_3term(cycle_t *c, const double *diag, double *Qb, const double *ab,
                                  double *bb, double *dbar, double *gbar)
{
    const int n = c->n;
    const size_t cs = 2 * (size_t)n;
    double *wb = (double *)malloc(cs * sizeof(double));
    double *Hw = (double *)malloc(cs * sizeof(double));
    for (int j = c->jb; j >= 1; --j) {
        const double *qn = c->Q + cs * (j + 1);
        const double *q = c->Q + cs * j;
        const double *qp = c->Q + cs * (j - 1);
        const double bj = c->beta[j], bp = c->beta[j - 1], aj = c->alpha[j];
        memset(wb, 0, cs * sizeof(double));
        if (bj > 0.0) {
            const double *qnb = Qb + cs * (j + 1);
            double rq = 0.0;
            for (size_t i = 0; i < cs; ++i) rq += qnb[i] * qn[i];
            double btot = bb[j] - rq / bj;
            for (size_t i = 0; i < cs; ++i) wb[i] = qnb[i] / bj + btot * qn[i];
        }
        /* w' = w - alpha q_j - beta_{j-1} q_{j-1} */
        double r2 = 0.0, r3 = 0.0;
        for (size_t i = 0; i < cs; ++i) { r2 += q[i] * wb[i]; r3 += qp[i] * wb[i]; }
        double atot = ab[j] - r2;
        if (j >= 2) bb[j - 1] += -r3;
        double *qjb = Qb + cs * j, *qpb = Qb + cs * (j - 1);
        for (size_t i = 0; i < cs; ++i) {
            double wj = bj * qn[i] + aj * q[i] + bp * qp[i];     /* H q_j */
            qjb[i] += -aj * wb[i] + atot * wj;
            if (j >= 2) qpb[i] += -bp * wb[i];
            wb[i] += atot * q[i];                               /* wbar */
        }
        orc_spmv(n, c->rowptr, c->col, c->val, c->g, diag, wb, Hw);
        for (size_t i = 0; i < cs; ++i) qjb[i] += Hw[i];
        for (int i = 0; i < n; ++i) dbar[i] += wb[2 * i] * q[2 * i] + wb[2 * i + 1] * q[2 * i + 1];
        jv(c, wb, Hw);                                          /* J wbar */
        double gs = 0.0;
        for (size_t i = 0; i < cs; ++i) gs += Hw[i] * q[i];
        *gbar += gs;
    }
    free(wb); free(Hw);
}

static int g_adjoint_mode = 1;   /* 0: exact CGS reverse, 1: three-term reverse */
void orc_set_adjoint_mode(int mode) { g_adjoint_mode = mode; }

/* reverse of one cycle: pbar (adjoint of the normalised output) ->
 * pbar (adjoint of the input); accumulates dbar[i] and *gbar */
static void cycle_reverse(cycle_t *c, const double *diag, double *pbar, double *dbar, double *gbar)
{
    const int n = c->n, m = c->m, nt = m + 2;
    const size_t cs = 2 * (size_t)n;
    const double inv = 1.0 / c->nrm;
    /* psi'' = psi'/|psi'| */
    double dot = 0.0;
    for (size_t i = 0; i < cs; ++i) dot += c->psi1[i] * inv * pbar[i];
    double *p1b = (double *)malloc(cs * sizeof(double));
    for (size_t i = 0; i < cs; ++i) p1b[i] = (pbar[i] - c->psi1[i] * inv * dot) * inv;
    /* combine psi' = sum_k y_k q_k */
    double *Qb = (double *)calloc(cs * (m + 2), sizeof(double));
    double *Eb = (double *)calloc((size_t)2 * nt * nt, sizeof(double));
    for (int k = 1; k <= m; ++k) {
        const double *q = c->Q + cs * k;
        double *qb = Qb + cs * k;
        double yr = c->y[2 * (k - 1)], yi = c->y[2 * (k - 1) + 1];
        double sr = 0.0, si = 0.0;
        for (int i = 0; i < n; ++i) {
            sr += q[2 * i] * p1b[2 * i] + q[2 * i + 1] * p1b[2 * i + 1];
            si += q[2 * i] * p1b[2 * i + 1] - q[2 * i + 1] * p1b[2 * i];
            qb[2 * i] += yr * p1b[2 * i] + yi * p1b[2 * i + 1];
            qb[2 * i + 1] += yr * p1b[2 * i + 1] - yi * p1b[2 * i];
        }
        Eb[2 * (k - 1)] = sr;            /* E[k-1, 0] */
        Eb[2 * (k - 1) + 1] = si;
    }
    /* Pade */
    double *Ab = (double *)calloc((size_t)2 * nt * nt, sizeof(double));
    pade_reverse(&c->tape, Eb, Ab);
    /* A = i cim Tbar (cim = -theta): Tbar_bar = Re(conj(Ab) * i cim) = cim * Im(Ab) */
    const double cim = -c->theta;
    double *ab = (double *)calloc(m + 2, sizeof(double));
    double *bb = (double *)calloc(m + 2, sizeof(double));
    for (int i = 0; i < m; ++i) {
        ab[i + 1] = cim * Ab[2 * (i + i * nt) + 1];
        if (i + 1 < m)
            bb[i + 1] = cim * (Ab[2 * (i + (i + 1) * nt) + 1] + Ab[2 * ((i + 1) + i * nt) + 1]);
    }
    if (g_adjoint_mode == 1) {
        lanczos_reverse_3term(c, diag, Qb, ab, bb, dbar, gbar);
        memcpy(pbar, Qb + cs, cs * sizeof(double));
        free(p1b); free(Qb); free(Eb); free(Ab); free(ab); free(bb);
        return;
    }
    /* Lanczos, j = jb .. 1 */
    double *wb = (double *)malloc(cs * sizeof(double));
    double *jq = (double *)malloc(cs * sizeof(double));
    double *hb = (double *)malloc(sizeof(double) * 2 * (m + 2));
    for (int j = c->jb; j >= 1; --j) {
        const double *qn = c->Q + cs * (j + 1);
        const double *q = c->Q + cs * j;
        const double *w = c->W + cs * j;
        const double bj = c->beta[j];
        /* q_{j+1} = w'/beta, beta = |w'| ; beta_j enters Tbar when j < m */
        memset(wb, 0, cs * sizeof(double));
        if (bj > 0.0) {
            const double *qnb = Qb + cs * (j + 1);
            double rq = 0.0;
            for (size_t i = 0; i < cs; ++i) rq += qnb[i] * qn[i];
            double btot = bb[j] - rq / bj;
            for (size_t i = 0; i < cs; ++i) wb[i] = qnb[i] / bj + btot * qn[i];
        }
        /* w' = w - sum_k h_k q_k */
        for (int k = 1; k <= j; ++k) {
            const double *qk = c->Q + cs * k;
            double *qkb = Qb + cs * k;
            double hr = HJ(c, j, k)[0], hi = HJ(c, j, k)[1];
            double sr = 0.0, si = 0.0;
            for (int i = 0; i < n; ++i) {
                sr += qk[2 * i] * wb[2 * i] + qk[2 * i + 1] * wb[2 * i + 1];
                si += qk[2 * i] * wb[2 * i + 1] - qk[2 * i + 1] * wb[2 * i];
                qkb[2 * i] -= hr * wb[2 * i] + hi * wb[2 * i + 1];
                qkb[2 * i + 1] -= hr * wb[2 * i + 1] - hi * wb[2 * i];
            }
            hb[2 * k] = -sr;
            hb[2 * k + 1] = -si;
        }
        hb[2 * j] += ab[j];          /* alpha_j = Re h_j */
        /* h_k = q_k^H w:  wbar += q_k hbar_k ; qbar_k += w conj(hbar_k) */
        double *wtot = jq;
        memcpy(wtot, wb, cs * sizeof(double));
        for (int k = 1; k <= j; ++k) {
            const double *qk = c->Q + cs * k;
            double *qkb = Qb + cs * k;
            double hr = hb[2 * k], hi = hb[2 * k + 1];
            for (int i = 0; i < n; ++i) {
                wtot[2 * i] += qk[2 * i] * hr - qk[2 * i + 1] * hi;
                wtot[2 * i + 1] += qk[2 * i] * hi + qk[2 * i + 1] * hr;
                qkb[2 * i] += w[2 * i] * hr + w[2 * i + 1] * hi;
                qkb[2 * i + 1] += w[2 * i + 1] * hr - w[2 * i] * hi;
            }
        }
        /* w = (diag + g J) q_j */
        double *qjb = Qb + cs * j;
        double *Hw = wb;   /* reuse */
        orc_spmv(n, c->rowptr, c->col, c->val, c->g, diag, wtot, Hw);
        for (size_t i = 0; i < cs; ++i) qjb[i] += Hw[i];
        for (int i = 0; i < n; ++i)
            dbar[i] += wtot[2 * i] * q[2 * i] + wtot[2 * i + 1] * q[2 * i + 1];
        double *Jq = Hw;
        jv(c, q, Jq);
        double gs = 0.0;
        for (size_t i = 0; i < cs; ++i) gs += wtot[i] * Jq[i];
        *gbar += gs;
    }
    memcpy(pbar, Qb + cs, cs * sizeof(double));
    free(p1b); free(Qb); free(Eb); free(Ab); free(ab); free(bb); free(wb); free(jq); free(hb);
}

double orc_adjoint(const orc_params *p, const orc_site_params *sp,
                   const int32_t *rowptr, const int32_t *col, const double *val,
                   const double *psi0, uint64_t traj, int64_t nsteps,
                   const int32_t *group, int ngroups, const double *w_pop,
                   const double *bra, double w_spec, const double *corr_target,
                   const double *corr_weight, int ckpt, double *grad_e, double *grad_g)
{
    const int n = p->n, m = p->m, K = p->krylov_substeps;
    const size_t cs = 2 * (size_t)n;
    if (ckpt < 1) ckpt = 1;
    const int64_t nck = (nsteps + ckpt - 1) / ckpt + 1;
    double *ck_psi = (double *)malloc(cs * nck * sizeof(double));
    double *ck_de = (double *)malloc((size_t)n * nck * sizeof(double));
    double *e0 = (double *)malloc(sizeof(double) * n);
    double *de = (double *)malloc(sizeof(double) * n);
    double *diag = (double *)malloc(sizeof(double) * n);
    double *psi = (double *)malloc(cs * sizeof(double));
    double *corr = (double *)calloc(2 * (size_t)(nsteps + 1), sizeof(double));
    cycle_t cyc;
    cycle_alloc(&cyc, n, m);
    cyc.rowptr = rowptr; cyc.col = col; cyc.val = val; cyc.g = p->g;
    cyc.theta = (p->dt_fs / (double)K) / p->hbar;
    const double tol = p->breakdown_tol;
    double loss = 0.0;
    /* ---------------- forward with checkpoints ---------------- */
    orc_init_disorder(p, sp, traj, e0, de);
    memcpy(psi, psi0, cs * sizeof(double));
    for (int64_t step = 0; step <= nsteps; ++step) {
        if (step % ckpt == 0) {
            memcpy(ck_psi + cs * (step / ckpt), psi, cs * sizeof(double));
            memcpy(ck_de + (size_t)n * (step / ckpt), de, n * sizeof(double));
        }
        if (bra) {
            double cr = 0.0, ci = 0.0;
            for (int i = 0; i < n; ++i) {
                cr += bra[2 * i] * psi[2 * i] + bra[2 * i + 1] * psi[2 * i + 1];
                ci += bra[2 * i] * psi[2 * i + 1] - bra[2 * i + 1] * psi[2 * i];
            }
            corr[2 * step] = cr;
            corr[2 * step + 1] = ci;
        }
        if (step == nsteps) break;
        orc_noise_step(p, sp, traj, step, de);
        for (int i = 0; i < n; ++i) diag[i] = de[i] + e0[i];
        for (int k = 0; k < K; ++k) {
            if (cycle_forward(&cyc, diag, psi, tol) != 0) { loss = NAN; goto out; }
            for (size_t i = 0; i < cs; ++i) psi[i] = cyc.psi1[i] / cyc.nrm;
        }
    }
    /* loss and its adjoint at T */
    double *pbar = (double *)calloc(cs, sizeof(double));
    for (int i = 0; i < n; ++i) {
        int gi = group ? group[i] : -1;
        if (gi >= 0 && gi < ngroups && w_pop) {
            loss += w_pop[gi] * (psi[2 * i] * psi[2 * i] + psi[2 * i + 1] * psi[2 * i + 1]);
            pbar[2 * i] += 2.0 * w_pop[gi] * psi[2 * i];
            pbar[2 * i + 1] += 2.0 * w_pop[gi] * psi[2 * i + 1];
        }
    }
    double *cbar = (double *)calloc(2 * (size_t)(nsteps + 1), sizeof(double));
    if (bra && w_spec != 0.0) {
        for (int64_t s = 0; s <= nsteps; ++s) {
            double dr = corr[2 * s] - (corr_target ? corr_target[2 * s] : 0.0);
            double di = corr[2 * s + 1] - (corr_target ? corr_target[2 * s + 1] : 0.0);
            const double ws = w_spec * (corr_weight ? corr_weight[s] : 1.0);
            loss += ws * (dr * dr + di * di);
            cbar[2 * s] = 2.0 * ws * dr;
            cbar[2 * s + 1] = 2.0 * ws * di;
        }
    }
    memset(grad_e, 0, sizeof(double) * n);
    double gg = 0.0;
    double *dbar = (double *)calloc(n, sizeof(double));
    double *seg_psi = (double *)malloc(cs * (size_t)(ckpt * K + 1) * sizeof(double));
    double *seg_diag = (double *)malloc(sizeof(double) * n * (size_t)ckpt);
    /* C(T) term */
    if (bra && w_spec != 0.0)
        for (int i = 0; i < n; ++i) {
            pbar[2 * i] += bra[2 * i] * cbar[2 * nsteps] - bra[2 * i + 1] * cbar[2 * nsteps + 1];
            pbar[2 * i + 1] += bra[2 * i] * cbar[2 * nsteps + 1] + bra[2 * i + 1] * cbar[2 * nsteps];
        }
    /* ---------------- reverse over segments ---------------- */
    for (int64_t seg = (nsteps - 1) / ckpt; seg >= 0; --seg) {
        int64_t s0 = seg * ckpt, s1 = s0 + ckpt < nsteps ? s0 + ckpt : nsteps;
        /* recompute the segment: noise replay from the checkpoint */
        memcpy(psi, ck_psi + cs * seg, cs * sizeof(double));
        memcpy(de, ck_de + (size_t)n * seg, n * sizeof(double));
        memcpy(seg_psi, psi, cs * sizeof(double));
        for (int64_t step = s0; step < s1; ++step) {
            orc_noise_step(p, sp, traj, step, de);
            double *dg = seg_diag + (size_t)n * (step - s0);
            for (int i = 0; i < n; ++i) dg[i] = de[i] + e0[i];
            for (int k = 0; k < K; ++k) {
                cycle_forward(&cyc, dg, psi, tol);
                for (size_t i = 0; i < cs; ++i) psi[i] = cyc.psi1[i] / cyc.nrm;
                memcpy(seg_psi + cs * ((step - s0) * K + k + 1), psi, cs * sizeof(double));
            }
        }
        for (int64_t step = s1 - 1; step >= s0; --step) {
            double *dg = seg_diag + (size_t)n * (step - s0);
            memset(dbar, 0, sizeof(double) * n);
            for (int k = K - 1; k >= 0; --k) {
                const double *in = seg_psi + cs * ((step - s0) * K + k);
                cycle_forward(&cyc, dg, in, tol);
                cycle_reverse(&cyc, dg, pbar, dbar, &gg);
            }
            for (int i = 0; i < n; ++i) grad_e[i] += dbar[i];
            /* C(step) term */
            if (bra && w_spec != 0.0)
                for (int i = 0; i < n; ++i) {
                    pbar[2 * i] += bra[2 * i] * cbar[2 * step] - bra[2 * i + 1] * cbar[2 * step + 1];
                    pbar[2 * i + 1] += bra[2 * i] * cbar[2 * step + 1] + bra[2 * i + 1] * cbar[2 * step];
                }
        }
    }
    *grad_g = gg;
    free(pbar); free(cbar); free(dbar); free(seg_psi); free(seg_diag);
out:
    cycle_free(&cyc);
    free(ck_psi); free(ck_de); free(e0); free(de); free(diag); free(psi); free(corr);
    return loss;
}
