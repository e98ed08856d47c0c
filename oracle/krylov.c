/*
 * krylov.c -- oracle restatement of the propagation step.  TEST INFRASTRUCTURE.
 *
 * One time step (propagate.cu:199-260):
 *   noise update of the diagonal       systemClass.cu:1940-1971, gpuSysHelper.cu:73-128
 *   Lanczos from psi                   lanczosClass.cu:230-353, 395-531
 *   augmented Tbar                     gpuLanczosHelper.cu:14-60
 *   Pade(6,6) scaling and squaring     padeClass.cu:204-487, 584-745
 *   psi = Q_m exp(A)[0:m,0]            propStep.cu:31-44
 *   renormalise                        systemClass.cu:1769-1829
 *   Expokit error estimate             padeClass.cu:749-782, systemClass.cu:2055-2085
 * with the deviations the engine makes on purpose: Hermitian (conjugated)
 * alpha instead of cublasZdotu (lanczosClass.cu:421), the true infinity norm
 * sum |A_ij| instead of sum |Re A_ij| (gpuPadeHelper.cu:25), optional full
 * reorthogonalisation, a breakdown guard, and K Krylov substeps per dt.
 */
#include "dq_oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

double orc_hbar_ref(void)
{
    /* cppUtility.cpp:19-22 evaluated in the same order */
    double pi = atan(1) * 4;
    double lightspeed = 299792548.;
    return (1 / 100.) * (1e15) * 1. / (lightspeed * 2 * pi);
}

void orc_pade_coeffs(int p, double *c)
{
    /* padeClass.cu:161-168 */
    c[0] = 1;
    double fp = (double)p;
    for (int k = 1; k <= p; ++k) {
        double fk = (double)k;
        c[k] = c[k - 1] * (fp + 1. - fk) / ((2 * fp + 1. - fk) * fk);
    }
}

void orc_spmv(int n, const int32_t *rowptr, const int32_t *col,
              const double *val, double g, const double *diag,
              const double *x, double *y)
{
    for (int i = 0; i < n; ++i) {
        double ar = 0.0, ai = 0.0;
        for (int32_t p = rowptr[i]; p < rowptr[i + 1]; ++p) {
            int j = col[p];
            ar += val[p] * x[2 * j];
            ai += val[p] * x[2 * j + 1];
        }
        y[2 * i] = diag[i] * x[2 * i] + g * ar;
        y[2 * i + 1] = diag[i] * x[2 * i + 1] + g * ai;
    }
}

/* <a|b> = sum conj(a_i) b_i */
static void cdot(int n, const double *a, const double *b, double *re, double *im)
{
    double sr = 0.0, si = 0.0;
    for (int i = 0; i < n; ++i) {
        sr += a[2 * i] * b[2 * i] + a[2 * i + 1] * b[2 * i + 1];
        si += a[2 * i] * b[2 * i + 1] - a[2 * i + 1] * b[2 * i];
    }
    *re = sr;
    *im = si;
}

static double cnorm(int n, const double *a)
{
    double s = 0.0;
    for (int i = 0; i < 2 * n; ++i) s += a[i] * a[i];
    return sqrt(s);
}

int orc_lanczos(int n, const int32_t *rowptr, const int32_t *col,
                const double *val, double g, const double *diag,
                const double *psi, const orc_lanczos_opts *o,
                double *Q, double *alpha, double *beta, double *nrm_hq)
{
    const int m = o->m;
    const size_t cs = 2 * (size_t)n;
    double *w = (double *)malloc(cs * sizeof(double));
    /* projection coefficients h^(j)_k (complex) and Gram entries G[l][j] */
    double *hh = (double *)calloc((size_t)2 * (m + 2) * (m + 2), sizeof(double));
    double *G = (double *)calloc((size_t)2 * (m + 2) * (m + 2), sizeof(double));
#define HH(j, k) (hh + 2 * ((size_t)(j) * (m + 2) + (k)))
#define GG(l, j) (G + 2 * ((size_t)(l) * (m + 2) + (j)))
    memset(Q, 0, cs * (m + 2) * sizeof(double));
    memcpy(Q + cs, psi, cs * sizeof(double));
    for (int j = 0; j <= m; ++j) { alpha[j] = 0.0; beta[j] = 0.0; }
    GG(1, 1)[0] = 1.0;
    int broke = 0;
    for (int j = 1; j <= m; ++j) {
        double *qj = Q + cs * j;
        orc_spmv(n, rowptr, col, val, g, diag, qj, w);
        if (o->reorth == ORC_REORTH_NONE) {
            double ar, ai;
            cdot(n, qj, w, &ar, &ai);
            alpha[j] = ar;
            const double *qm = Q + cs * (j - 1);
            for (size_t i = 0; i < cs; ++i) w[i] = w[i] - ar * qj[i] - beta[j - 1] * qm[i];
        } else {
            double *h = HH(j, 0);
            cdot(n, qj, w, &h[2 * j], &h[2 * j + 1]);
            if (o->reorth == ORC_REORTH_CGS) {
                for (int k = 1; k < j; ++k)
                    cdot(n, Q + cs * k, w, &h[2 * k], &h[2 * k + 1]);
            } else {
                /* q_k^H H q_j = (H q_k)^H q_j with H q_k = sum_l h^(k)_l q_l
                 * + beta_k q_{k+1}: projections from the Gram row of q_j. */
                for (int k = 1; k < j; ++k) {
                    double sr = beta[k] * GG(k + 1, j)[0];
                    double si = beta[k] * GG(k + 1, j)[1];
                    for (int l = 1; l <= k; ++l) {
                        const double *a = HH(k, l), *b = GG(l, j);
                        sr += a[0] * b[0] + a[1] * b[1];
                        si += a[0] * b[1] - a[1] * b[0];
                    }
                    h[2 * k] = sr;
                    h[2 * k + 1] = si;
                }
            }
            alpha[j] = h[2 * j];
            for (int k = 1; k <= j; ++k) {
                const double *qk = Q + cs * k;
                double hr = h[2 * k], hi = h[2 * k + 1];
                for (int i = 0; i < n; ++i) {
                    double xr = qk[2 * i], xi = qk[2 * i + 1];
                    w[2 * i] -= hr * xr - hi * xi;
                    w[2 * i + 1] -= hr * xi + hi * xr;
                }
            }
        }
        double b = cnorm(n, w);
        double *qn = Q + cs * (j + 1);
        if (!(b > o->breakdown_tol * (fabs(alpha[j]) + beta[j - 1]))) {
            broke = 1;
            beta[j] = 0.0;
            break;                          /* q_{j+1..} stay zero */
        }
        beta[j] = b;
        double s = 1.0 / b;
        for (size_t i = 0; i < cs; ++i) qn[i] = w[i] * s;
        if (o->reorth == ORC_REORTH_GRAM) {
            for (int l = 1; l <= j; ++l)
                cdot(n, Q + cs * l, qn, &GG(l, j + 1)[0], &GG(l, j + 1)[1]);
            GG(j + 1, j + 1)[0] = 1.0;
            GG(j + 1, j + 1)[1] = 0.0;
        }
    }
    if (!broke) {
        orc_spmv(n, rowptr, col, val, g, diag, Q + cs * (m + 1), w);
        *nrm_hq = cnorm(n, w);
    } else {
        *nrm_hq = 0.0;
    }
#undef HH
#undef GG
    free(w);
    free(hh);
    free(G);
    return broke;
}

void orc_build_tbar_aug(int m, int aug, const double *alpha, const double *beta, double *tb)
{
    const int n = m + aug;
    memset(tb, 0, sizeof(double) * n * n);
    for (int i = 0; i < m; ++i) {
        tb[i + i * n] = alpha[i + 1];
        if (i + 1 < m) {
            tb[i + (i + 1) * n] = beta[i + 1];
            tb[(i + 1) + i * n] = beta[i + 1];
        }
    }
    tb[0 + m * n] = 1.0;
    for (int i = m; i + 1 < n; ++i) tb[i + (i + 1) * n] = 1.0;
}

void orc_build_tbar(int m, const double *alpha, const double *beta, double *tb)
{
    const int n = m + 2;
    memset(tb, 0, sizeof(double) * n * n);
    for (int i = 0; i < m; ++i) {
        tb[i + i * n] = alpha[i + 1];
        if (i + 1 < m) {
            tb[i + (i + 1) * n] = beta[i + 1];
            tb[(i + 1) + i * n] = beta[i + 1];
        }
    }
    tb[0 + m * n] = 1.0;
    tb[m + (m + 1) * n] = 1.0;
}

/* C = A B, complex column-major n x n */
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

/* Solve D X = B (complex, n x n, nrhs = n) in place in B by Gaussian
 * elimination with partial pivoting on |re|+|im| (culaDeviceZgesv role,
 * padeClass.cu:714-745).  D is destroyed. */
static int csolve(int n, double *D, double *B)
{
    for (int k = 0; k < n; ++k) {
        int p = k;
        double best = -1.0;
        for (int i = k; i < n; ++i) {
            double v = fabs(D[2 * (i + k * n)]) + fabs(D[2 * (i + k * n) + 1]);
            if (v > best) { best = v; p = i; }
        }
        if (!(best > 0.0)) return -1;
        if (p != k) {
            for (int j = 0; j < n; ++j) {
                double *a = D + 2 * (k + j * n), *b = D + 2 * (p + j * n);
                double t0 = a[0], t1 = a[1]; a[0] = b[0]; a[1] = b[1]; b[0] = t0; b[1] = t1;
                a = B + 2 * (k + j * n); b = B + 2 * (p + j * n);
                t0 = a[0]; t1 = a[1]; a[0] = b[0]; a[1] = b[1]; b[0] = t0; b[1] = t1;
            }
        }
        double pr = D[2 * (k + k * n)], pi = D[2 * (k + k * n) + 1];
        double den = pr * pr + pi * pi;
        double ir = pr / den, ii = -pi / den;       /* 1/pivot */
        for (int i = k + 1; i < n; ++i) {
            double *a = D + 2 * (i + k * n);
            double lr = a[0] * ir - a[1] * ii, li = a[0] * ii + a[1] * ir;
            a[0] = lr; a[1] = li;
            for (int j = k + 1; j < n; ++j) {
                double *d = D + 2 * (i + j * n);
                const double *u = D + 2 * (k + j * n);
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
    for (int j = 0; j < n; ++j) {
        for (int i = n - 1; i >= 0; --i) {
            double sr = B[2 * (i + j * n)], si = B[2 * (i + j * n) + 1];
            for (int k = i + 1; k < n; ++k) {
                const double *u = D + 2 * (i + k * n), *x = B + 2 * (k + j * n);
                sr -= u[0] * x[0] - u[1] * x[1];
                si -= u[0] * x[1] + u[1] * x[0];
            }
            double pr = D[2 * (i + i * n)], pi = D[2 * (i + i * n) + 1];
            double den = pr * pr + pi * pi;
            B[2 * (i + j * n)] = (sr * pr + si * pi) / den;
            B[2 * (i + j * n) + 1] = (si * pr - sr * pi) / den;
        }
    }
    return 0;
}

int orc_pade_expm(int n, const double *A_in, double *E, int *s_out)
{
    const size_t sz = (size_t)2 * n * n;
    double c[7];
    orc_pade_coeffs(6, c);
    double *A = (double *)malloc(sz * sizeof(double) * 8);
    double *A2 = A + sz, *A4 = A2 + sz, *A6 = A4 + sz;
    double *T1 = A6 + sz, *Num = T1 + sz, *Den = Num + sz, *Tmp = Den + sz;
    memcpy(A, A_in, sz * sizeof(double));
    /* true infinity norm (doScalingStep, padeClass.cu:584-649, with
     * |A_ij| instead of |Re A_ij|) */
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
    *s_out = s;
    if (s >= 50) { free(A); return -1; }
    double sc = ldexp(1.0, -s);
    for (size_t i = 0; i < sz; ++i) A[i] *= sc;
    cmm(n, A, A, A2);
    cmm(n, A2, A2, A4);
    cmm(n, A2, A4, A6);
    /* U = A (c1 I + c3 A^2 + c5 A^4) */
    for (size_t i = 0; i < sz; ++i) T1[i] = c[3] * A2[i] + c[5] * A4[i];
    for (int i = 0; i < n; ++i) T1[2 * (i + i * n)] += c[1];
    cmm(n, A, T1, Num);
    /* V = c0 I + c2 A^2 + c4 A^4 + c6 A^6 ; Den = V - U */
    for (size_t i = 0; i < sz; ++i) Den[i] = c[2] * A2[i] + c[4] * A4[i] + c[6] * A6[i] - Num[i];
    for (int i = 0; i < n; ++i) Den[2 * (i + i * n)] += c[0];
    memcpy(Tmp, Num, sz * sizeof(double));
    if (csolve(n, Den, Tmp) != 0) { free(A); return -2; }
    /* E = I + 2 X, then square s times (padeClass.cu:451-476, 656-706) */
    for (size_t i = 0; i < sz; ++i) E[i] = 2.0 * Tmp[i];
    for (int i = 0; i < n; ++i) E[2 * (i + i * n)] += 1.0;
    for (int k = 0; k < s; ++k) {
        cmm(n, E, E, Tmp);
        memcpy(E, Tmp, sz * sizeof(double));
    }
    free(A);
    return 0;
}

static double expokit_err_ld(int m, int n, double beta_m, double nrm_hq, const double *E)
{
    double hm = fabs(beta_m);
    const double *e1 = E + 2 * ((m - 1) + m * n);
    const double *e2 = E + 2 * ((m - 1) + (m + 1) * n);
    double err1 = hm * sqrt(e1[0] * e1[0] + e1[1] * e1[1]);
    double err2 = hm * nrm_hq * sqrt(e2[0] * e2[0] + e2[1] * e2[1]);
    if (err1 > err2) return err2 / (1 - err2 / err1);
    return err1;
}

double orc_expokit_err(int m, double beta_m, double nrm_hq, const double *E)
{
    return expokit_err_ld(m, m + 2, beta_m, nrm_hq, E);
}

/* ------------------------------------------------------------------ */
void orc_init_disorder(const orc_params *p, const orc_site_params *sp,
                       uint64_t traj, double *e0, double *de)
{
    for (int i = 0; i < p->n; ++i) {
        double base = sp->e_site ? sp->e_site[i] : 0.0;
        if (p->sigma_static > 0.0) {
            double z = orc_normal(p->seed, traj, (uint64_t)i, 0, 1, 0);
            e0[i] = p->sigma_static * z + base;
        } else {
            e0[i] = base;
        }
        if (sp->e_shift) e0[i] = e0[i] + sp->e_shift[i];
        /* the stationary OU start (systemClass.cu:1036); the other models
         * set the offset at every step */
        double sig = (sp->noise_model == ORC_NOISE_OU && sp->ou_sigma) ? sp->ou_sigma[i] : 0.0;
        de[i] = sig * orc_normal(p->seed, traj, (uint64_t)i, 0, 2, 0);
    }
}

double orc_zz_phase(uint64_t seed, uint64_t traj, uint64_t site, uint64_t k)
{
    uint64_t ctr[4] = {site, k >> 2, 5, 0}, key[2] = {seed, traj}, out[4];
    orc_philox4x64_10(ctr, key, out);
    return (double)(out[k & 3] >> 11) * (1.0 / 9007199254740992.0);
}

double orc_zz_cos(double u, double w2pi, double t)
{
    double x = u + w2pi * t;
    x = x - floor(x);
    if (x >= 1.0) x = 0.0;
    double s, c;
    orc_sincos_2pi(x, &s, &c);
    return c;
}

void orc_zz_oscillators(int nosc, const double *omega_cm, const double *J, double domega,
                        double temperature, double *w2pi, double *coeff)
{
    const double kB = 0.695035, conv = 5308.84, pi = 3.14159265358979323846;
    for (int k = 0; k < nosc; ++k) {
        double G = J[k] / (1 - exp(-1. / (kB * temperature) * omega_cm[k]));
        coeff[k] = sqrt(2.) * sqrt(G * domega / pi);
        w2pi[k] = omega_cm[k] / conv / (2 * pi);
    }
}

void orc_noise_step(const orc_params *p, const orc_site_params *sp,
                    uint64_t traj, int64_t step, double *de)
{
    switch (sp->noise_model) {
    case ORC_NOISE_OU:
        orc_ou_step(p, sp, traj, step, de);
        break;
    case ORC_NOISE_HSR:
        for (int i = 0; i < p->n; ++i) {
            double sig = sp->hsr_sigma ? sp->hsr_sigma[i] : 0.0;
            de[i] = sig * orc_normal(p->seed, traj, (uint64_t)i, (uint64_t)step, 4, 0);
        }
        break;
    case ORC_NOISE_ZZ: {
        const double t = (double)step * p->dt_fs;
        for (int i = 0; i < p->n; ++i) {
            double acc = 0.0;
            for (int k = 0; k < sp->zz_nosc; ++k)
                acc = acc + sp->zz_coeff[k] * orc_zz_cos(orc_zz_phase(p->seed, traj, (uint64_t)i, (uint64_t)k),
                                                         sp->zz_w2pi[k], t);
            de[i] = acc;
        }
        break;
    }
    default:
        for (int i = 0; i < p->n; ++i) de[i] = 0.0;
    }
}

void orc_ou_step(const orc_params *p, const orc_site_params *sp,
                 uint64_t traj, int64_t step, double *de)
{
    const int S = p->ou_substeps;
    double subdt = p->dt_fs / (double)S;
    double sqrt_subdt = sqrt(subdt);
    for (int i = 0; i < p->n; ++i) {
        if (!sp->ou_sigma || !sp->ou_tau || sp->ou_sigma[i] == 0.0) continue;
        double invTau = 1. / sp->ou_tau[i];
        double coeff = sp->ou_sigma[i] * sqrt(2. / sp->ou_tau[i]);
        double cur = de[i];
        for (int s = 0; s < S; ++s) {
            double xi = orc_normal(p->seed, traj, (uint64_t)i, (uint64_t)step, 3, (uint64_t)s);
            cur = cur - cur * invTau * subdt + coeff * xi * sqrt_subdt;
        }
        de[i] = cur;
    }
}

/* one Krylov cycle psi <- Q_m exp(-i tau Tbar/hbar)[0:m,0]; returns
 * clamped error estimate, or a negative value on failure */
static double krylov_cycle(int n, const int32_t *rowptr, const int32_t *col,
                           const double *val, double g, const double *diag,
                           const orc_lanczos_opts *lo, double tau, double hbar,
                           double *psi, double *Q, double *work)
{
    const int m = lo->m, nt = m + 2;
    double *alpha = work, *beta = alpha + (m + 2);
    double *tb = beta + (m + 2);
    double *A = tb + nt * nt;
    double *E = A + 2 * nt * nt;
    double nrm_hq;
    orc_lanczos(n, rowptr, col, val, g, diag, psi, lo, Q, alpha, beta, &nrm_hq);
    orc_build_tbar(m, alpha, beta, tb);
    double cim = (-1. / hbar) * tau;         /* propagate.cu:91-92 */
    for (int i = 0; i < nt * nt; ++i) { A[2 * i] = 0.0; A[2 * i + 1] = cim * tb[i]; }
    int s;
    if (orc_pade_expm(nt, A, E, &s) != 0) return -1.0;
    const size_t cs = 2 * (size_t)n;
    for (int i = 0; i < n; ++i) {
        double sr = 0.0, si = 0.0;
        for (int k = 1; k <= m; ++k) {
            const double *q = Q + cs * k + 2 * i;
            double yr = E[2 * (k - 1)], yi = E[2 * (k - 1) + 1];
            sr += yr * q[0] - yi * q[1];
            si += yr * q[1] + yi * q[0];
        }
        psi[2 * i] = sr;
        psi[2 * i + 1] = si;
    }
    double err = orc_expokit_err(m, beta[m], nrm_hq, E);
    if (err < 1.110223e-16) err = 1.110223e-16;   /* includeFiles.cu:10 */
    return err;
}

void orc_tk_coeffs(int p, double tau, double hbar, double *coef)
{
    /* h_expCoeff = -i dt / hbar (propagate.cu:89-93); powComplex by repeated
     * complex products from 1 (cppUtility.cpp:52-70), divided by k! */
    const double cr = 0.0, ci = (-1. / hbar) * tau;
    double zr = 1.0, zi = 0.0, fact = 1.0;
    for (int k = 0; k < p; ++k) {
        if (k > 0) {
            const double r = zr * cr - zi * ci, i = zr * ci + zi * cr;
            zr = r;
            zi = i;
            fact *= (double)k;
        }
        coef[2 * k] = zr / fact;
        coef[2 * k + 1] = zi / fact;
    }
}

double orc_tk_cycle(int n, const int32_t *rowptr, const int32_t *col, const double *val, double g,
                    const double *diag, const orc_params *p, double tau, double *psi)
{
    const int P = p->tk_p, m = p->tk_m, nt = m + P;
    const size_t cs = 2 * (size_t)n;
    double *t = (double *)malloc(cs * sizeof(double) * 3);
    double *u = t + cs, *y = u + cs;
    double *Q = (double *)malloc(cs * (m + 2) * sizeof(double));
    double *coef = (double *)malloc(sizeof(double) * 2 * P);
    double *alpha = (double *)malloc(sizeof(double) * (2 * (m + 2) + nt * nt * 5));
    double *beta = alpha + (m + 2), *tb = beta + (m + 2), *A = tb + nt * nt, *E = A + 2 * nt * nt;
    orc_tk_coeffs(P, tau, p->hbar, coef);
    /* Taylor part: t = sum_{k<p} c^k A^k psi / k!, u = A^p psi
     * (taylorKrylovClass.cu:226-322) */
    memcpy(t, psi, cs * sizeof(double));
    memcpy(u, psi, cs * sizeof(double));
    for (int k = 1; k <= P; ++k) {
        orc_spmv(n, rowptr, col, val, g, diag, u, y);
        if (k < P) {
            const double ar = coef[2 * k], ai = coef[2 * k + 1];
            for (int i = 0; i < n; ++i) {
                t[2 * i] += ar * y[2 * i] - ai * y[2 * i + 1];
                t[2 * i + 1] += ar * y[2 * i + 1] + ai * y[2 * i];
            }
        }
        memcpy(u, y, cs * sizeof(double));
    }
    /* Krylov part on the auxiliary matrix from A^p psi / |A^p psi|
     * (taylorKrylovClass.cu:345-418, calcAndStoreNormAndScale) */
    const double nu = cnorm(n, u);
    const double inv = nu > 0.0 ? 1.0 / nu : 0.0;
    for (size_t i = 0; i < cs; ++i) u[i] *= inv;
    orc_lanczos_opts lo = {m, p->reorth, p->breakdown_tol};
    double nrm_hq;
    orc_lanczos(n, p->tk_rowptr, p->tk_col, p->tk_val, 1.0, p->tk_diag, u, &lo, Q, alpha, beta, &nrm_hq);
    orc_build_tbar_aug(m, P, alpha, beta, tb);
    const double cim = (-1. / p->hbar) * tau;
    for (int i = 0; i < nt * nt; ++i) { A[2 * i] = 0.0; A[2 * i + 1] = cim * tb[i]; }
    int s;
    double err = -1.0;
    if (orc_pade_expm(nt, A, E, &s) == 0) {
        /* psi = t + |A^p psi| Q_m E[0:m, m+p-1] (taylorKrylovClass.cu:372-418) */
        const double *col_last = E + 2 * (size_t)nt * (nt - 1);
        for (int i = 0; i < n; ++i) {
            double sr = 0.0, si = 0.0;
            for (int k = 1; k <= m; ++k) {
                const double *q = Q + cs * k + 2 * i;
                const double yr = col_last[2 * (k - 1)], yi = col_last[2 * (k - 1) + 1];
                sr += yr * q[0] - yi * q[1];
                si += yr * q[1] + yi * q[0];
            }
            psi[2 * i] = t[2 * i] + nu * sr;
            psi[2 * i + 1] = t[2 * i + 1] + nu * si;
        }
        err = expokit_err_ld(m, nt, beta[m], nrm_hq, E);
    }
    free(t); free(Q); free(coef); free(alpha);
    return err;
}

static void record(int n, const double *psi, const int32_t *group, int ngroups, double *out)
{
    for (int g = 0; g < ngroups; ++g) out[g] = 0.0;
    for (int i = 0; i < n; ++i) {
        int g = group ? group[i] : i;
        if (g < 0 || g >= ngroups) continue;
        out[g] += psi[2 * i] * psi[2 * i] + psi[2 * i + 1] * psi[2 * i + 1];
    }
}

/* participation ratio (systemClass.cu:2209-2257): the reference sums
 * (re^2 + im^2)^2 over the sites and stores the reciprocal */
double orc_participation(int n, const double *psi)
{
    double s = 0.0;
    for (int i = 0; i < n; ++i) {
        double pp = psi[2 * i] * psi[2 * i] + psi[2 * i + 1] * psi[2 * i + 1];
        s += pp * pp;
    }
    return 1.0 / s;
}

int orc_run(const orc_params *p, const orc_site_params *sp,
            const int32_t *rowptr, const int32_t *col, const double *val,
            const double *psi0, uint64_t traj, int64_t nsteps,
            int obs_every, const int32_t *group, int ngroups,
            const double *bra, double *pops, double *corr, double *partic,
            double *psi_out, double *err_sum)
{
    const int n = p->n, m = p->m, nt = m + 2;
    const size_t cs = 2 * (size_t)n;
    double *psi = (double *)malloc(cs * sizeof(double));
    double *Q = (double *)malloc(cs * (m + 2) * sizeof(double));
    double *e0 = (double *)malloc(sizeof(double) * n * 3);
    double *de = e0 + n, *diag = de + n;
    double *work = (double *)malloc(sizeof(double) * (2 * (m + 2) + nt * nt * 5));
    orc_lanczos_opts lo = {m, p->reorth, p->breakdown_tol};
    double esum = 0.0;
    int rc = 0;
    memcpy(psi, psi0, cs * sizeof(double));
    /* the initial state is normalised (systemClass.cu:1383-1481,
     * normalizeHostCol): psi / sqrt(sum |psi_i|^2), summed site by site */
    {
        double s2 = 0.0;
        for (int i = 0; i < n; ++i) s2 += psi[2 * i] * psi[2 * i] + psi[2 * i + 1] * psi[2 * i + 1];
        if (!(s2 > 0.0)) { rc = -2; goto done; }
        double inv = 1.0 / sqrt(s2);
        for (size_t i = 0; i < cs; ++i) psi[i] *= inv;
    }
    orc_init_disorder(p, sp, traj, e0, de);
    double tau = p->dt_fs / (double)p->krylov_substeps;
    if (pops && obs_every > 0) record(n, psi, group, ngroups, pops);
    if (corr && bra) cdot(n, bra, psi, &corr[0], &corr[1]);
    if (partic) partic[0] = orc_participation(n, psi);
    for (int64_t step = 0; step < nsteps; ++step) {
        orc_noise_step(p, sp, traj, step, de);
        for (int i = 0; i < n; ++i) diag[i] = de[i] + e0[i];
        for (int k = 0; k < p->krylov_substeps; ++k) {
            double err;
            if (p->tk_p > 0) {
                err = orc_tk_cycle(n, rowptr, col, val, p->g, diag, p, tau, psi);
                if (err < 0.0) { rc = -1; goto done; }
            } else {
                err = krylov_cycle(n, rowptr, col, val, p->g, diag, &lo, tau,
                                   p->hbar, psi, Q, work);
                if (err < 0.0) { rc = -1; goto done; }
            }
            double nr = cnorm(n, psi);
            if (p->tk_p > 0) {
                /* getTaylorKrylovError adds |1 - |psi||, then the clamp
                 * (taylorKrylovClass.cu:426-440, systemClass.cu:2071-2077) */
                err = err + fabs(1. - nr);
                if (err < 1.110223e-16) err = 1.110223e-16;
            }
            esum += err;
            double inv = 1.0 / nr;
            for (size_t i = 0; i < cs; ++i) psi[i] *= inv;
        }
        if (pops && obs_every > 0 && (step + 1) % obs_every == 0)
            record(n, psi, group, ngroups, pops + (size_t)ngroups * ((step + 1) / obs_every));
        if (corr && bra) cdot(n, bra, psi, &corr[2 * (step + 1)], &corr[2 * (step + 1) + 1]);
        if (partic) partic[step + 1] = orc_participation(n, psi);
    }
done:
    if (psi_out) memcpy(psi_out, psi, cs * sizeof(double));
    if (err_sum) *err_sum = esum;
    free(psi); free(Q); free(e0); free(work);
    return rc;
}

int orc_run_batch(const orc_params *p, const orc_site_params *sp,
                  const int32_t *rowptr, const int32_t *col, const double *val,
                  const double *psi0, const uint64_t *traj_ids, int nb,
                  const int32_t *n_active, int64_t nsteps, int obs_every,
                  const int32_t *group, int ngroups, const double *bra,
                  double *pops, double *corr, double *partic, double *err_sum, int nthreads)
{
    const int n = p->n;
    const size_t nobs = obs_every > 0 ? (size_t)(nsteps / obs_every + 1) : 0;
    int fail = 0;
#ifdef _OPENMP
    if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads) reduction(| : fail)
#endif
    for (int b = 0; b < nb; ++b) {
        int nb_act = n_active ? n_active[b] : n;
        orc_params pb = *p;
        pb.n = nb_act;
        /* leading nb_act x nb_act block of the coupling matrix */
        int32_t *rp = (int32_t *)malloc(sizeof(int32_t) * (nb_act + 1));
        int64_t cnt = 0;
        rp[0] = 0;
        for (int i = 0; i < nb_act; ++i) {
            for (int32_t q = rowptr[i]; q < rowptr[i + 1]; ++q) if (col[q] < nb_act) ++cnt;
            rp[i + 1] = (int32_t)cnt;
        }
        int32_t *cl = (int32_t *)malloc(sizeof(int32_t) * (cnt + 1));
        double *vl = (double *)malloc(sizeof(double) * (cnt + 1));
        cnt = 0;
        for (int i = 0; i < nb_act; ++i)
            for (int32_t q = rowptr[i]; q < rowptr[i + 1]; ++q)
                if (col[q] < nb_act) { cl[cnt] = col[q]; vl[cnt] = val[q]; ++cnt; }
        /* the auxiliary Taylor-Krylov matrix, cut to the same leading block */
        int32_t *trp = NULL, *tcl = NULL;
        double *tvl = NULL;
        if (p->tk_p > 0) {
            trp = (int32_t *)malloc(sizeof(int32_t) * (nb_act + 1));
            int64_t tc = 0;
            trp[0] = 0;
            for (int i = 0; i < nb_act; ++i) {
                for (int32_t q = p->tk_rowptr[i]; q < p->tk_rowptr[i + 1]; ++q) if (p->tk_col[q] < nb_act) ++tc;
                trp[i + 1] = (int32_t)tc;
            }
            tcl = (int32_t *)malloc(sizeof(int32_t) * (tc + 1));
            tvl = (double *)malloc(sizeof(double) * (tc + 1));
            tc = 0;
            for (int i = 0; i < nb_act; ++i)
                for (int32_t q = p->tk_rowptr[i]; q < p->tk_rowptr[i + 1]; ++q)
                    if (p->tk_col[q] < nb_act) { tcl[tc] = p->tk_col[q]; tvl[tc] = p->tk_val[q]; ++tc; }
            pb.tk_rowptr = trp;
            pb.tk_col = tcl;
            pb.tk_val = tvl;
        }
        int r = orc_run(&pb, sp, rp, cl, vl, psi0 + (size_t)b * 2 * n, traj_ids[b], nsteps,
                        obs_every, group, ngroups, bra,
                        pops ? pops + (size_t)b * nobs * ngroups : NULL,
                        corr ? corr + (size_t)b * 2 * (nsteps + 1) : NULL,
                        partic ? partic + (size_t)b * (nsteps + 1) : NULL,
                        NULL, err_sum ? err_sum + b : NULL);
        if (r != 0) fail = 1;
        free(rp); free(cl); free(vl);
        free(trp); free(tcl); free(tvl);
    }
    return fail ? -1 : 0;
}

/* Substep control restated (the engine's dq_system_set_krylov_control): the
 * batch advances in chunks of `chunk` steps with K Krylov cycles per dt; a
 * chunk is kept when K times the largest clamped Expokit estimate of any
 * cycle of any trajectory is at most tol, and is otherwise redone from its
 * start with 2K cycles (K <= kmax).  K starts at p->krylov_substeps.
 * Populations every obs_every steps (pops[b][rec][g]), final states
 * psi_out[b][2n], error sums err_sum[b], the K of every step k_of_step[step]
 * and the number of redone chunks. */
int orc_run_controlled(const orc_params *p, const orc_site_params *sp,
                       const int32_t *rowptr, const int32_t *col, const double *val,
                       const double *psi0, const uint64_t *traj, int nb, int64_t nsteps,
                       double tol, int kmax, int chunk, int obs_every, const int32_t *group, int ngroups,
                       double *pops, double *psi_out, double *err_sum, int32_t *k_of_step, int64_t *rejects)
{
    const int n = p->n, m = p->m, nt = m + 2;
    const size_t cs = 2 * (size_t)n;
    const size_t nobs = obs_every > 0 ? (size_t)(nsteps / obs_every + 1) : 0;
    double *psi = (double *)malloc(sizeof(double) * cs * nb * 2);
    double *psi_ck = psi + cs * nb;
    double *e0 = (double *)malloc(sizeof(double) * n * nb * 3);
    double *de = e0 + (size_t)n * nb, *de_ck = de + (size_t)n * nb;
    double *esum = (double *)calloc(2 * (size_t)nb, sizeof(double));
    double *esum_ck = esum + nb;
    double *diag = (double *)malloc(sizeof(double) * n);
    double *Q = (double *)malloc(sizeof(double) * cs * (m + 2));
    double *work = (double *)malloc(sizeof(double) * (2 * (m + 2) + nt * nt * 5));
    orc_lanczos_opts lo = {m, p->reorth, p->breakdown_tol};
    int rc = 0, K = p->krylov_substeps;
    *rejects = 0;
    for (int b = 0; b < nb; ++b) {
        double *pb = psi + cs * b;
        memcpy(pb, psi0 + cs * b, cs * sizeof(double));
        double s2 = 0.0;
        for (int i = 0; i < n; ++i) s2 += pb[2 * i] * pb[2 * i] + pb[2 * i + 1] * pb[2 * i + 1];
        double inv = 1.0 / sqrt(s2);
        for (size_t i = 0; i < cs; ++i) pb[i] *= inv;
        orc_init_disorder(p, sp, traj[b], e0 + (size_t)n * b, de + (size_t)n * b);
        if (pops && obs_every > 0) record(n, pb, group, ngroups, pops + (size_t)b * nobs * ngroups);
    }
    int64_t done = 0;
    while (done < nsteps) {
        const int64_t c = nsteps - done < chunk ? nsteps - done : chunk;
        memcpy(psi_ck, psi, sizeof(double) * cs * nb);
        memcpy(de_ck, de, sizeof(double) * n * nb);
        memcpy(esum_ck, esum, sizeof(double) * nb);
        const double tau = p->dt_fs / (double)K;
        double emax = 0.0;
        for (int b = 0; b < nb; ++b) {
            double *pb = psi + cs * b, *db = de + (size_t)n * b, *eb = e0 + (size_t)n * b;
            for (int64_t step = done; step < done + c; ++step) {
                orc_noise_step(p, sp, traj[b], step, db);
                for (int i = 0; i < n; ++i) diag[i] = db[i] + eb[i];
                for (int k = 0; k < K; ++k) {
                    double err = krylov_cycle(n, rowptr, col, val, p->g, diag, &lo, tau, p->hbar, pb, Q, work);
                    if (err < 0.0) { rc = -1; goto done_label; }
                    if (err > emax) emax = err;
                    esum[b] += err;
                    double inv = 1.0 / cnorm(n, pb);
                    for (size_t i = 0; i < cs; ++i) pb[i] *= inv;
                }
                if (pops && obs_every > 0 && (step + 1) % obs_every == 0)
                    record(n, pb, group, ngroups, pops + ((size_t)b * nobs + (step + 1) / obs_every) * ngroups);
            }
        }
        if (!(emax * K <= tol) && 2 * K <= kmax) {
            memcpy(psi, psi_ck, sizeof(double) * cs * nb);
            memcpy(de, de_ck, sizeof(double) * n * nb);
            memcpy(esum, esum_ck, sizeof(double) * nb);
            K *= 2;
            ++*rejects;
            continue;
        }
        for (int64_t step = done; step < done + c; ++step)
            if (k_of_step) k_of_step[step] = K;
        done += c;
    }
done_label:
    if (psi_out) memcpy(psi_out, psi, sizeof(double) * cs * nb);
    if (err_sum) memcpy(err_sum, esum, sizeof(double) * nb);
    free(psi); free(e0); free(esum); free(diag); free(Q); free(work);
    return rc;
}
