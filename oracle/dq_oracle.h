/*
 * dq_oracle.h -- FP64 CPU restatement of DynaQuant's stochastic exciton
 * propagation method.  TEST INFRASTRUCTURE ONLY: this library is the
 * checker for the HIP engine in dynaquant_amd/.  Only tests/, the smoke
 * check and the CPU-baseline leg of bench.py load it; the product never
 * links, imports or calls it.
 *
 * Every routine cites the reference file:line whose behaviour it restates.
 * Complex vectors are interleaved doubles (re, im).  Dense small matrices
 * are column-major, like the reference's denseMatrixClass.
 *
 * Pinning (see tests/test_oracle_*.py):
 *   Philox4x64-10          numpy.random.Philox known answers (tests/golden)
 *   normals (Box-Muller)   libm log/sin/cos to 4 ulp, moments
 *   OU update              stationary variance, lag autocorrelation
 *   tube couplings         independent numpy dipole formula, nnz counts
 *   SpMV                   dense numpy product
 *   Lanczos                Krylov relation, orthonormality, N<=m exactness
 *   Pade expm              scipy.linalg.expm, scalar R66, 2x2 rotation,
 *                          Pade coefficients of the reference's run log
 *   Krylov step            numpy eigh dense propagator, Rabi closed form
 *   FFT/spectrum           numpy.fft, brute DFT, single-site closed form
 *   adjoint                central finite differences of orc_run
 *   participation ratio    closed forms (k equal weights -> k, two weights)
 *   CD vectors (pyoracle)  numpy.cross, orthogonality to R and mu
 */
#ifndef DQ_ORACLE_H
#define DQ_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- constants ------------------------------------------------------- */
/* hbar in cm^-1 * fs with the reference's lightspeed 299792548 m/s
 * (cppUtility.cpp:10-35; run log out:76 prints 5308.84). */
double orc_hbar_ref(void);
/* Pade(p,p) numerator coefficients c_0..c_p (padeClass.cu:161-168). */
void orc_pade_coeffs(int p, double *c);

/* ---- counter-based RNG (replaces the MRG32k3a states of
 * gpuSysHelper.cu:10-26 and Boost mt19937 of systemClass.cu:759-804) ---- */
void orc_philox4x64_10(const uint64_t ctr[4], const uint64_t key[2], uint64_t out[4]);
double orc_log(double x);                 /* x in (0, inf), own polynomial */
double orc_sqrt(double x);                /* deterministic, no hw sqrt */
void orc_sincos_2pi(double u, double *s, double *c); /* sin/cos(2 pi u), u in [0,1) */
void orc_normal4(uint64_t seed, uint64_t traj, uint64_t c0, uint64_t c1,
                 uint64_t c2, uint64_t c3, double z[4]);
/* one standard normal: stream 1 = static disorder, 2 = OU initial value,
 * 3 = OU increment; `sub` selects the OU substep. */
double orc_normal(uint64_t seed, uint64_t traj, uint64_t site, uint64_t step,
                  uint64_t stream, uint64_t sub);

/* ---- tube geometry and couplings (brute force, all pairs) ------------ */
/* walls: n_walls walls, wall w has rings[w] rings of `per_ring` sites at
 * radius[w] nm.  order 0 = wall-major (i = wall_off + per_ring*r + k),
 * order 1 = ring-major (i = r*(n_walls*per_ring) + w*per_ring + k, only
 * valid when all walls have the same ring count).  Returns N. */
int orc_tube_sites(int n_walls, const int *rings, const double *radius,
                   int per_ring, double ring_dz, double mu_debye,
                   double tilt_deg, int order,
                   double *pos, double *dip, int *wall_of_site);
/* J_ij = kappa*(mu_i.mu_j - 3(mu_i.r)(mu_j.r))/r^3, r < cutoff, i != j.
 * Two calls: with val==NULL it only fills rowptr and returns nnz. */
int64_t orc_couplings_csr(int n, const double *pos, const double *dip,
                          double cutoff_nm, double kappa,
                          int32_t *rowptr, int32_t *col, double *val);

/* ---- Krylov units ---------------------------------------------------- */
/* y = (diag + g*J) x, complex x/y (lanczosClass.cu:618-661) */
void orc_spmv(int n, const int32_t *rowptr, const int32_t *col,
              const double *val, double g, const double *diag,
              const double *x, double *y);

enum { ORC_REORTH_NONE = 0, ORC_REORTH_CGS = 1, ORC_REORTH_GRAM = 2 };

typedef struct {
    int m;               /* Krylov depth */
    int reorth;          /* ORC_REORTH_* */
    double breakdown_tol;
} orc_lanczos_opts;

/* Hermitian Lanczos (lanczosClass.cu:395-531 with the conjugated dot).
 * Q: (m+2) columns of n complex, column 0 unused (zero), q_1 = psi.
 * alpha[1..m], beta[1..m] (index 0 unused).  *nrm_hq = ||H q_{m+1}||
 * (lanczosClass.cu:536-570).  Returns number of breakdowns detected (0/1). */
int orc_lanczos(int n, const int32_t *rowptr, const int32_t *col,
                const double *val, double g, const double *diag,
                const double *psi, const orc_lanczos_opts *o,
                double *Q, double *alpha, double *beta, double *nrm_hq);

/* Augmented Tbar (m+2)^2 real, column-major (gpuLanczosHelper.cu:14-60). */
void orc_build_tbar(int m, const double *alpha, const double *beta, double *tbar);

/* Pade(6,6) with scaling and squaring of a general complex n x n matrix
 * (padeClass.cu:204-487, with the true infinity norm).  A, E complex
 * column-major.  Returns 0 or -1 when s >= 50 (padeClass.cu:618-623). */
int orc_pade_expm(int n, const double *A, double *E, int *s_out);

/* Expokit error estimate (padeClass.cu:749-782, systemClass.cu:2055-2085). */
double orc_expokit_err(int m, double beta_m, double nrm_hq, const double *E);
/* Augmented tridiagonal of size m + aug (gpuLanczosHelper.cu:14-60): T in the
 * leading m x m block, 1 at (0, m) and at (i, i + 1) for m <= i < m + aug - 1;
 * aug = 2 is orc_build_tbar.  Column-major. */
void orc_build_tbar_aug(int m, int aug, const double *alpha, const double *beta, double *tb);
/* Taylor coefficients c^k / k!, k = 0..p-1, of c = -i tau / hbar
 * (taylorKrylovClass.cu:200-214), interleaved complex */
void orc_tk_coeffs(int p, double tau, double hbar, double *coef);

/* ---- trajectory ------------------------------------------------------ */
typedef struct {
    int n;
    int m;
    int krylov_substeps;   /* K */
    int ou_substeps;       /* S (ka_substeps) */
    int reorth;
    double breakdown_tol;
    double dt_fs;
    double hbar;           /* cm^-1 fs */
    double g;              /* coupling scale */
    double sigma_static;   /* cm^-1 */
    uint64_t seed;
    /* Taylor-Krylov propagation (taylorKrylovClass.cu:133-434) when tk_p > 0:
     * order p of the Taylor part, m of the auxiliary Krylov part, and the
     * auxiliary matrix (off-diagonal CSR + diagonal), read from tkmatfile */
    int tk_p, tk_m;
    const int32_t *tk_rowptr, *tk_col;
    const double *tk_val, *tk_diag;
} orc_params;

/* site-energy noise models (systemClass.cu:1881-2013): Kubo-Anderson /
 * Ornstein-Uhlenbeck (SIM_TYPE_KUBO_ANDERSON), Haken-Strobl-Reineker white
 * noise (SIM_TYPE_HSR), Zhong-Zhao real noise (SIM_TYPE_ZZ_REAL_NOISE), none
 * (SIM_TYPE_CONST_HAM) */
enum { ORC_NOISE_OU = 0, ORC_NOISE_HSR = 1, ORC_NOISE_ZZ = 2, ORC_NOISE_NONE = 3 };

typedef struct {
    const double *ou_tau;    /* [n] fs */
    const double *ou_sigma;  /* [n] cm^-1 */
    const double *e_site;    /* [n] site energies before disorder */
    const double *e_shift;   /* [n] or NULL: added to e0 (adjoint direction) */
    int noise_model;         /* ORC_NOISE_* */
    int zz_nosc;             /* ZZ oscillators per site */
    const double *hsr_sigma; /* [n] cm^-1, HSR standard deviations */
    const double *zz_w2pi;   /* [zz_nosc] omega / (2 pi) in 1/fs */
    const double *zz_coeff;  /* [zz_nosc] amplitudes in cm^-1 */
} orc_site_params;

/* Run one trajectory `traj`.  Outputs (any may be NULL):
 * pops[(nsteps/obs_every+1)*ngroups], corr[2*(nsteps+1)] = <bra|psi(n)>,
 * partic[nsteps+1] = 1 / sum_i |psi_i(n)|^4 (systemClass.cu:2209-2257),
 * psi_out[2n], err_sum (sum of clamped Expokit estimates).
 * Returns 0 on success, <0 on error. */
int orc_run(const orc_params *p, const orc_site_params *sp,
            const int32_t *rowptr, const int32_t *col, const double *val,
            const double *psi0, uint64_t traj, int64_t nsteps,
            int obs_every, const int32_t *group, int ngroups,
            const double *bra, double *pops, double *corr, double *partic,
            double *psi_out, double *err_sum);
/* 1 / sum_i |psi_i|^4 of psi[2n] */
double orc_participation(int n, const double *psi);

/* B trajectories traj_ids[b], OpenMP over b (threads = nthreads, 0 = all).
 * pops laid out [b][obs][group], corr [b][n+1][2], partic [b][n+1]. */
int orc_run_batch(const orc_params *p, const orc_site_params *sp,
                  const int32_t *rowptr, const int32_t *col, const double *val,
                  const double *psi0, const uint64_t *traj_ids, int nb,
                  const int32_t *n_active, int64_t nsteps, int obs_every,
                  const int32_t *group, int ngroups, const double *bra,
                  double *pops, double *corr, double *partic, double *err_sum, int nthreads);
/* substep control from the Expokit estimate, a batch advanced chunk by chunk */
int orc_run_controlled(const orc_params *p, const orc_site_params *sp,
                       const int32_t *rowptr, const int32_t *col, const double *val,
                       const double *psi0, const uint64_t *traj, int nb, int64_t nsteps,
                       double tol, int kmax, int chunk, int obs_every, const int32_t *group, int ngroups,
                       double *pops, double *psi_out, double *err_sum, int32_t *k_of_step, int64_t *rejects);

/* One Taylor-Krylov cycle on psi (taylorKrylovClass.cu:217-434), A = the
 * Hamiltonian (val, g, diag), the auxiliary matrix from p.  Returns the
 * Expokit estimate of the auxiliary decomposition (unclamped) or < 0. */
double orc_tk_cycle(int n, const int32_t *rowptr, const int32_t *col, const double *val, double g,
                    const double *diag, const orc_params *p, double tau, double *psi);

/* Static site energies e0 and initial OU offsets of trajectory traj
 * (systemClass.cu:759-804, 1028-1036). */
void orc_init_disorder(const orc_params *p, const orc_site_params *sp,
                       uint64_t traj, double *e0, double *de);
/* One OU update of step `step` (gpuSysHelper.cu:73-128). */
void orc_ou_step(const orc_params *p, const orc_site_params *sp,
                 uint64_t traj, int64_t step, double *de);
/* The noise offset of step `step` for the selected model: OU update of de
 * (above); HSR de_i = sigma_i Z(i, step, stream 4) (gpuSysHelper.cu:32-66);
 * ZZ de_i = sum_k c_k cos(2 pi (u_ik + w_k t)), t = step dt, u_ik uniform
 * phases of Philox stream 5 fixed per trajectory (gpuSysHelper.cu:134-177,
 * systemClass.cu:1075-1243); none: 0. */
void orc_noise_step(const orc_params *p, const orc_site_params *sp,
                    uint64_t traj, int64_t step, double *de);
/* uniform phase u in [0,1) of ZZ oscillator k of `site` (Philox stream 5) */
double orc_zz_phase(uint64_t seed, uint64_t traj, uint64_t site, uint64_t k);
/* cos(2 pi (u + w t)) with the argument reduced to [0,1) and the oracle's
 * own sin/cos of 2 pi x */
double orc_zz_cos(double u, double w2pi, double t);
/* ZZ amplitudes from a spectral density J(omega) sampled every domega
 * (systemClass.cu:1118-1145): G = J / (1 - exp(-omega/(kB T))),
 * c = sqrt(2) sqrt(G domega / pi), kB = 0.695035 cm^-1/K; omega/(2 pi) in
 * 1/fs with the reference's conversion 5308.84 cm^-1 fs (includeFiles.cu:16). */
void orc_zz_oscillators(int nosc, const double *omega_cm, const double *J, double domega,
                        double temperature, double *w2pi, double *coeff);

/* ---- spectra --------------------------------------------------------- */
/* in-place radix-2 FFT, sign=+1 inverse-direction (unnormalised) */
int orc_fft(int n, double *x, int sign);
void orc_dft(int n, const double *x, double *y, int sign);
/* Linear absorption from an averaged correlation function corr[0..T]
 * (spectraClass.cu:575-750): window w(n) = exp(-0.5((n+1)dt/sig)^2) for
 * n < T, zero-pad to nfft, inverse-direction FFT, scale mu2/T.
 * omega[k] in cm^-1 (k wraps above nfft/2), spec complex [2*nfft]. */
int orc_spectrum(int T, const double *corr, double dt_fs, double sigma_fs,
                 double mu2, int nfft, double *omega, double *spec);

/* ---- adjoint --------------------------------------------------------- */
/* Loss = sum_g w_pop[g] * P_g(T) + w_spec * sum_n u(n) |C(n) - C_target(n)|^2
 * on one trajectory (u = corr_weight, NULL: 1).  Gradient w.r.t. the n site energies and g.
 * Exact reverse mode through the discrete noise/Lanczos/Pade map with
 * checkpoints every `ckpt` steps.  Returns loss. */
double orc_adjoint(const orc_params *p, const orc_site_params *sp,
                   const int32_t *rowptr, const int32_t *col, const double *val,
                   const double *psi0, uint64_t traj, int64_t nsteps,
                   const int32_t *group, int ngroups, const double *w_pop,
                   const double *bra, double w_spec, const double *corr_target,
                   const double *corr_weight, int ckpt, double *grad_e, double *grad_g);

/* 0: reverse of the full-projection Lanczos as computed (O(m^2) vector ops);
 * 1 (default): reverse of the equivalent three-term recurrence on the
 * stored basis (O(m)), the form the engine uses. */
void orc_set_adjoint_mode(int mode);

#ifdef __cplusplus
}
#endif
#endif
