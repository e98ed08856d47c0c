/*
 * dynaquant.h -- C ABI of the MI355X (gfx950) stochastic exciton-dynamics
 * engine.  Plain C types only: interleaved (re, im) doubles for complex
 * vectors, int32 indices, int64 counts, opaque handles.
 *
 * The entry points replace the seams of the reference's one-executable
 * driver (propagate.cu:199-260):
 *   systemClass::parseInFile / readInHam     -> dq_system_create,
 *                                               dq_system_set_hamiltonian_csr
 *   systemClass::prepForKA                   -> dq_system_set_ou_params
 *   systemClass::resetStaticDisorder         -> dq_system_reset_disorder
 *   systemClass::setInitState                -> dq_system_set_state
 *   systemClass::updateHamiltonian           -> dq_update_noise
 *   lanczosClass::doLanczos + padeClass::doPade + propagateSystem
 *     + renormalizeAndStoreNorm + calculateErrorEstimate
 *                                            -> dq_propagate_step, dq_run
 *   popPrintClass::printOneLoopAllPops       -> dq_get_populations
 *   spectraClass::calcCorrFunc               -> dq_get_corr
 *   spectraClass::prepForCD / prepInitState  -> dq_cd_vectors
 *   systemClass::calcInvPatricRatio          -> dq_system_set_participation,
 *                                               dq_get_participation
 *   taylorKrylovClass (PROP_MODE_TAYLOR_LANCZOS)
 *                                            -> dq_system_set_taylor_krylov
 *   spectraClass::singleSpectraCalc          -> dq_spectrum
 * Every call returns a dq_status; the message of the last failure on the
 * calling thread is returned by dq_last_error().  Nothing prints, nothing
 * exits.
 */
#ifndef DYNAQUANT_H
#define DYNAQUANT_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum dq_status {
    DQ_OK = 0,
    DQ_ERR_INVALID_ARG = 1,
    DQ_ERR_OOM = 2,
    DQ_ERR_HIP = 3,
    DQ_ERR_BREAKDOWN = 4,        /* reserved: breakdowns are handled, not fatal */
    DQ_ERR_NONFINITE = 5,
    DQ_ERR_SCALING_OVERFLOW = 6, /* Pade scaling exponent s >= 50 (padeClass.cu:618-623) */
    DQ_ERR_NO_DEVICE = 7,
    DQ_ERR_STATE = 8             /* call out of order (e.g. run before set_state) */
} dq_status;

typedef enum dq_precision {
    DQ_PREC_FP64 = 0,   /* complex128 vectors, FP64 everything */
    DQ_PREC_MIXED = 1   /* complex64 Krylov vectors and SpMM; FP64 psi, alpha/beta,
                           reorthogonalisation sums, Pade */
} dq_precision;

/* Site-energy noise on the diagonal, updated before every step
 * (systemClass::updateHamiltonian, systemClass.cu:1881-2013). */
typedef enum dq_noise {
    DQ_NOISE_OU = 0,    /* Kubo-Anderson / Ornstein-Uhlenbeck, dq_system_set_ou_params
                           (SIM_TYPE_KUBO_ANDERSON) */
    DQ_NOISE_HSR = 1,   /* Haken-Strobl-Reineker white noise, dq_system_set_hsr_params
                           (SIM_TYPE_HSR) */
    DQ_NOISE_ZZ = 2,    /* Zhong-Zhao real noise, dq_system_set_zz_params
                           (SIM_TYPE_ZZ_REAL_NOISE) */
    DQ_NOISE_NONE = 3   /* constant Hamiltonian apart from static disorder
                           (SIM_TYPE_CONST_HAM) */
} dq_noise;

typedef struct dq_system dq_system;

/* infile.in keys of the reference in brackets (systemClass.cu:231-630) */
typedef struct dq_config {
    int32_t n_sites;          /* N [hamfile size]; ragged batches: the largest N */
    int32_t n_traj;           /* B trajectories in one batch; 1 = plain path */
    int32_t krylov_m;         /* [m] Krylov depth, 2..32 */
    int32_t krylov_substeps;  /* K Krylov cycles per dt; 1 = the reference */
    int32_t ou_substeps;      /* [ka_substeps] Euler-Maruyama substeps per dt */
    int32_t precision;        /* dq_precision */
    int32_t force_batched;    /* 1: run B=1 through the batched kernels */
    int32_t noise_model;      /* [simtype] dq_noise, 0 = OU */
    double dt_fs;             /* [dt] */
    double hbar_cm_fs;        /* <= 0: the reference's 5308.84 (cppUtility.cpp:10-35) */
    double coupling_scale;    /* g: H = diag + g * couplings */
    double sigma_static;      /* [sigmastatdis] static disorder; 0 = [staticdis]=0 */
    double breakdown_tol;     /* <= 0: 1e-10 (FP64) / 1e-5 (MIXED) */
    uint64_t seed;            /* Philox key word 0 */
} dq_config;

/* ---- lifetime -------------------------------------------------------- */
dq_status dq_system_create(const dq_config *cfg, dq_system **out);
dq_status dq_system_destroy(dq_system *sys);
const char *dq_last_error(void);
dq_status dq_sync(dq_system *sys);

/* ---- system set-up ---------------------------------------------------- */
/* Off-diagonal couplings (cm^-1, real, row-sorted CSR without the diagonal)
 * and site energies e_site[n] (replaces the COO file read and cusparse
 * coo2csr of sparseMatrixClasses.cu:74-267, 489-540; the diagonal index map
 * of systemClass.cu:808-870 is not needed: the diagonal is held apart). */
dq_status dq_system_set_hamiltonian_csr(dq_system *sys, int32_t n, const int32_t *rowptr,
                                        const int32_t *col, const double *val,
                                        const double *e_site);
/* Per-site OU relaxation time (fs) and standard deviation (cm^-1)
 * (ka_paramsfile lines "tau sigma", systemClass.cu:980-1071). */
dq_status dq_system_set_ou_params(dq_system *sys, const double *tau_fs, const double *sigma);
/* HSR: per-site standard deviation (cm^-1) of a fresh Gaussian offset every
 * step (hsr parameter file, systemClass.cu:910-975; gpuSysHelper.cu:32-66;
 * Philox stream 4 instead of the shared MRG32k3a states). */
dq_status dq_system_set_hsr_params(dq_system *sys, const double *sigma);
/* ZZ real noise: offset_i(t) = sum_k coeff[k] cos(omega_k t + phi_ik), t = step
 * dt, omega_k = omega_cm[k] / 5308.84 rad/fs (includeFiles.cu:16), phases
 * uniform in [0, 2 pi) per site and trajectory from Philox stream 5 (the
 * reference draws them with Boost per run, systemClass.cu:1075-1243;
 * gpuSysHelper.cu:134-177).  coeff from dq_zz_oscillators. */
dq_status dq_system_set_zz_params(dq_system *sys, int32_t nosc, const double *omega_cm,
                                  const double *coeff);
/* Amplitudes of a spectral density J(omega_k) sampled every domega cm^-1 at
 * temperature T (K): G = J / (1 - exp(-omega / (kB T))), coeff = sqrt(2 G domega
 * / pi), kB = 0.695035 cm^-1/K (systemClass.cu:1118-1145). */
dq_status dq_zz_oscillators(int32_t nosc, const double *omega_cm, const double *J, double domega,
                            double temperature_K, double *coeff);
/* Taylor-Krylov propagation (propmode PROP_MODE_TAYLOR_LANCZOS,
 * taylorKrylovClass.cu:133-434) instead of the Lanczos propagator: per Krylov
 * cycle of length tau,
 *   psi' = sum_{k<p} c^k A^k psi / k! + |A^p psi| Q exp(c Tbar_p)[0:m_aux, m_aux+p-1],
 * c = -i tau / hbar, A the trajectory's Hamiltonian (noise included), Q and
 * Tbar_p (the tridiagonal augmented by p, gpuLanczosHelper.cu:14-60) from
 * m_aux Lanczos steps of the constant auxiliary matrix (rowptr, col, val
 * off-diagonal CSR as in dq_system_set_hamiltonian_csr, diagonal diag[n]; the
 * reference's tkmatfile) on A^p psi / |A^p psi|.  The error estimate per
 * cycle is the auxiliary Expokit estimate + |1 - |psi'|| (clamped,
 * taylorKrylovClass.cu:426-440).  Needs p >= 2, 2 <= m_aux <= m of the
 * config, m_aux + p <= 34 (the reference's defaults: p = 10, m_aux = 24).
 * p = 0 switches back to the Lanczos propagator.  Not with dq_adjoint. */
dq_status dq_system_set_taylor_krylov(dq_system *sys, int32_t p, int32_t m_aux, const int32_t *rowptr,
                                      const int32_t *col, const double *val, const double *diag);
/* Philox key word 1 of each trajectory (NULL: 0..B-1) and its active site
 * count (NULL: all N).  A trajectory with n_active < N propagates under the
 * leading n_active x n_active block of H.  Clears observables and the state
 * (call dq_system_reset_disorder and dq_system_set_state again).  Internally
 * the trajectories are ordered by n_active so that rows no trajectory of a
 * slice uses are skipped; every output stays in the caller's order. */
dq_status dq_system_set_trajectories(dq_system *sys, const uint64_t *traj_ids,
                                     const int32_t *n_active);
/* Draw static disorder and the stationary OU start for every trajectory,
 * reset the step counter and the error sums (systemClass.cu:759-804,
 * 1028-1036, 1552-1608). */
dq_status dq_system_reset_disorder(dq_system *sys);
/* psi[b][i] interleaved complex, b < B, i < N; entries at i >= n_active[b]
 * are ignored (systemClass.cu:1325-1508) */
dq_status dq_system_set_state(dq_system *sys, const double *psi);
dq_status dq_system_get_state(dq_system *sys, double *psi);
/* Resume point of a run: the state as the next step will use it (a
 * normalisation still pending from dq_system_set_state is applied first),
 * the OU offsets noise[b][i] (cm^-1), the error sums err_sum[b] (may be NULL)
 * and the step counter.  dq_system_restore puts them back, bit for bit, into
 * a system set up the same way (Hamiltonian, noise parameters, trajectories,
 * dq_system_reset_disorder): the static disorder and the noise of later
 * steps follow from the Philox counters, so a run split at any step and
 * continued in another process equals the unsplit run bit for bit.  Records of
 * observables before the restored step are not part of the snapshot.  No
 * counterpart in the reference (its runs restart from the initial state);
 * this is the checkpoint/resume of a long ensemble run. */
dq_status dq_system_snapshot(dq_system *sys, double *psi, double *noise, double *err_sum, int64_t *step);
dq_status dq_system_restore(dq_system *sys, const double *psi, const double *noise, const double *err_sum,
                            int64_t step);
/* Observables: site i adds |psi_i|^2 to group[i] (< 0: none), recorded
 * every `every` steps from step 0 up to max_steps; when ngroups == N and
 * group[i] == i the per-site populations are stored directly
 * (popPrintClass.cu:442-472).  bra[i] (interleaved complex, may be NULL)
 * enables C(n) = <bra|psi(n)> every step (spectraClass.cu:509-571). */
dq_status dq_system_set_observables(dq_system *sys, const int32_t *group, int32_t ngroups,
                                    int32_t every, int64_t max_steps, const double *bra);
/* Several bras: C_b(n) = <bras[bra_of_traj[b]]|psi_b(n)>, bras[nbra][n]
 * interleaved complex, bra_of_traj[b] < nbra per caller trajectory.  The
 * reference runs one ensemble per dipole direction d = x, y, z, each with
 * psi(0) = bra = mu_d / |mu_d| (spectraClass.cu:437-461), and averages the
 * three spectra (spectraClass.cu:699-744); here the three directions run as
 * one batch and dq_spectrum_weighted forms the average.  Call after
 * dq_system_set_observables (which returns to its single bra). */
dq_status dq_system_set_bras(dq_system *sys, int32_t nbra, const double *bras, const int32_t *bra_of_traj);
/* Participation ratio PR_b(n) = 1 / sum_i |psi_b,i(n)|^4 of the normalised
 * state, every step n = 0..max_steps of the current observables
 * (systemClass.cu:2209-2257, called before each step at propagate.cu:208;
 * the reference prints the ensemble average as "invParticRat.dat",
 * systemClass.cu:1737-1751).  Call after dq_system_set_observables, which
 * switches it off; on = 0 switches it off. */
dq_status dq_system_set_participation(dq_system *sys, int32_t on);

/* ---- Hamiltonian generator -------------------------------------------- */
/* Cylindrical aggregate of point dipoles: wall w has rings[w] rings of
 * per_ring sites at radius_nm[w]; ring r sits at z = r * ring_dz_nm and odd
 * rings are rotated by half a site; dipoles of mu_debye are tilted tilt_deg
 * from the ring tangent towards the axis; J_ij = kappa (mu_i.mu_j -
 * 3 (mu_i.r)(mu_j.r)) / r^3 for r < cutoff_nm.  order 0 numbers sites wall by
 * wall, order 1 ring by ring across walls (then every shorter tube with the
 * same walls is a leading block).  The reference instead reads a
 * precomputed Hamiltonian file (sparseMatrixClasses.cu:74-267). */
typedef struct dq_tube_spec {
    int32_t n_walls;
    int32_t per_ring;
    int32_t order;
    int32_t rings[4];
    double radius_nm[4];
    double ring_dz_nm;
    double mu_debye;
    double tilt_deg;
    double cutoff_nm;
    double kappa;        /* cm^-1 nm^3 / D^2 */
} dq_tube_spec;
/* pos[3n], dip[3n] (may be NULL), wall[n] (may be NULL); *n_out = n */
dq_status dq_tube_sites(const dq_tube_spec *spec, double *pos, double *dip, int32_t *wall,
                        int32_t *n_out);
/* Exactly symmetric CSR of the couplings, sorted columns, no diagonal.
 * With col == NULL only rowptr[n+1] and *nnz_out are filled. */
dq_status dq_tube_couplings(const dq_tube_spec *spec, int32_t *rowptr, int32_t *col, double *val,
                            int64_t *nnz_out);

/* ---- propagation ------------------------------------------------------ */
/* Advance the OU noise of the current step, write the diagonal and move
 * the step counter on without propagating (systemClass.cu:1881-1971,
 * gpuSysHelper.cu:73-128).  A unit seam: dq_run includes the noise. */
dq_status dq_update_noise(dq_system *sys);
/* One dt for all trajectories: noise, K x (Lanczos, Pade, combine,
 * renormalise, Expokit error), observables. */
dq_status dq_propagate_step(dq_system *sys);
/* nsteps steps, then renormalise and record the final state; returns when
 * the device is done.  use_graph != 0 replays one captured step as a
 * hipGraph (one graph launch per dt). */
dq_status dq_run(dq_system *sys, int64_t nsteps, int32_t use_graph);
dq_status dq_get_step(dq_system *sys, int64_t *step);
/* Krylov substeps from the Expokit error estimate (padeClass.cu:749-782,
 * systemClass.cu:2055-2085; the reference only sums the estimate).  With
 * tol > 0, dq_run advances in chunks of chunk_steps steps; a chunk is kept
 * when K times the largest clamped estimate of any Krylov cycle of any
 * trajectory is at most tol (the error allowed per dt), and is otherwise
 * redone from its start with twice the substeps, K <= k_max.  K starts at
 * krylov_substeps of the config and never decreases (the noise is
 * stationary).  tol <= 0 switches the control off and restores the
 * configured K.  Lanczos propagator only. */
dq_status dq_system_set_krylov_control(dq_system *sys, double tol, int32_t k_max, int32_t chunk_steps);
/* current K and the number of chunks redone so far */
dq_status dq_get_krylov_control(dq_system *sys, int32_t *k_now, int64_t *rejected_chunks);

/* ---- results ---------------------------------------------------------- */
/* pops[b][k][g] for records k = 0..step/every (*nrec set to that count). */
dq_status dq_get_populations(dq_system *sys, double *pops, int64_t *nrec);
/* corr[b][n] interleaved complex, n = 0..step */
dq_status dq_get_corr(dq_system *sys, double *corr);
/* pr[b][n], n = 0..step (dq_system_set_participation) */
dq_status dq_get_participation(dq_system *sys, double *pr);
/* sum over cycles of the clamped Expokit estimate, per trajectory */
dq_status dq_get_error_sum(dq_system *sys, double *err);

/* ---- unit seams (for tests and tools) --------------------------------- */
/* Current diagonal d[b][i] = static energy + OU offset of every trajectory
 * (the values updateHamKuboAnderson writes into the CSR diagonal,
 * gpuSysHelper.cu:121-123); FP32 in MIXED precision, returned as double. */
dq_status dq_get_diagonal(dq_system *sys, double *diag);
/* y = H_b x for trajectory b with its current diagonal (lanczosClass.cu:618-661) */
dq_status dq_spmv(dq_system *sys, int32_t traj, const double *x, double *y);
/* exp(-i theta Tbar) for `count` augmented tridiagonals (alpha[c][m],
 * beta[c][m]), complex column-major E[c][(m+2)^2]; s[c] = scaling exponent.
 * Runs the engine's Pade kernel (padeClass.cu:204-487). */
dq_status dq_pade_tbar(int32_t m, int32_t count, const double *alpha, const double *beta,
                       double theta, double *E, int32_t *s);
/* Spectrum from the batch-averaged C(n) (spectraClass.cu:575-750):
 * window exp(-0.5((n+1)dt/sigma)^2) for n < T (sigma <= 0: none),
 * zero-pad to nfft (power of two), inverse-direction FFT on the GPU,
 * scale mu2/T.  mu2 is the reference's scaling constant: |mu_d|^2 for
 * absorption, (pi/2) 1e-7 |mu_d| |R x mu_d| for circular dichroism
 * (spectraClass.cu:641-648, dq_cd_vectors).  omega[nfft] in cm^-1,
 * spec[nfft] interleaved complex. */
dq_status dq_spectrum(dq_system *sys, double sigma_fs, double mu2, int32_t nfft,
                      double *omega, double *spec);
/* dq_spectrum of C(n) = sum_b weight[b] C_b(n) (weights per caller
 * trajectory) and scale 1/T: with weight[b] = |mu_d(b)|^2 / (3 B_d) over three
 * direction ensembles of B_d trajectories each it is the reference's
 * orientational average of the absorption spectrum (spectraClass.cu:699-744;
 * dq_spectrum is the case weight[b] = mu2 / B). */
dq_status dq_spectrum_weighted(dq_system *sys, double sigma_fs, const double *weight, int32_t nfft,
                               double *omega, double *spec);
/* No counterpart in the reference: dq_spectrum for callers that cannot
 * allocate device memory (code under HIP graph capture, serving loops,
 * harnesses that forbid hipMalloc).  The windowed C(n) and the FFT's second
 * buffer live in a fixed per-GPU workspace of 2 * 131072 complex doubles
 * (4 MiB, held until the process ends): nfft <= 131072, else
 * DQ_ERR_INVALID_ARG.  Same result as dq_spectrum, bit for bit.  Calls that
 * overlap on one GPU share the workspace and race. */
dq_status dq_spectrum_noalloc(dq_system *sys, double sigma_fs, double mu2, int32_t nfft,
                              double *omega, double *spec);
/* In-place FFTs of `batch` interleaved complex rows of length n (power of
 * two, <= 65536) on the GPU (replaces cufftExecZ2Z, spectraClass.cu:617). */
dq_status dq_fft(int32_t n, int32_t batch, double *data, int32_t sign);

/* ---- adjoint ---------------------------------------------------------- */
/* Gradient of the per-trajectory loss
 *   L_b = sum_g w_pop[g] P_{b,g}(T) + w_spec sum_{n=0..T} u(n) |C_b(n) - C_target(n)|^2
 * (groups and bra from dq_system_set_observables; w_pop, C_target and the
 * weights u = corr_weight[T+1] may be NULL, u = 1).  With u(n) =
 * nfft (mu2/T)^2 win(n)^2 the second term is the L2 distance of the
 * windowed, zero-padded spectra of dq_spectrum (Parseval).  The gradient is
 * taken over nsteps steps from the current state at step 0, with respect
 * to the site energies (grad_e[b][i]) and the coupling scale (grad_g[b]).
 * Exact reverse sweep of the discrete map: states are checkpointed every
 * `ckpt` steps, each segment is recomputed with the noise replayed from
 * its Philox stream, and the Lanczos part is reversed through its
 * three-term recurrence.  Batched FP64 systems only.  No counterpart in
 * the reference (it has no derivatives). */
dq_status dq_adjoint(dq_system *sys, int64_t nsteps, int32_t ckpt, const double *w_pop,
                     double w_spec, const double *corr_target, const double *corr_weight,
                     double *loss, double *grad_e, double *grad_g);

/* ---- timing ----------------------------------------------------------- */
/* Milliseconds of the last dq_run, measured with HIP events on the
 * system's stream. */
dq_status dq_last_run_ms(dq_system *sys, double *ms);
/* Kernel classes timed by the profiling mode. */
enum {
    DQ_K_PROLOGUE = 0, DQ_K_SPMM = 1, DQ_K_SCALARS = 2, DQ_K_UPDATE = 3, DQ_K_REDUCE = 4,
    DQ_K_SPMM_LAST = 5, DQ_K_PADE = 6, DQ_K_COMBINE = 7, DQ_K_NORM = 8, DQ_K_CLASSES = 9
};
/* on != 0: later dq_run calls launch eagerly and bracket every kernel with
 * HIP events on the system's stream; times accumulate per class until the
 * mode is switched on again.  ms[DQ_K_CLASSES], launches[DQ_K_CLASSES]. */
dq_status dq_set_profiling(dq_system *sys, int32_t on);
dq_status dq_get_kernel_times(dq_system *sys, double *ms, int64_t *launches);

/* Launch choices of the batched kernels.  No counterpart in the reference
 * (its csrmv/zgemv launch shapes are fixed inside cuSPARSE/cuBLAS); these are
 * the search space of dq_autotune.  Every choice gives the same per-trajectory
 * arithmetic; tile_rows only regroups the FP64 partial sums.  The plain
 * single-trajectory path has fixed shapes and ignores them. */
typedef struct dq_tuning {
    int32_t spmm_rows_per_wave;  /* 2, 4 or 8 rows whose column union one wavefront gathers;
                                    16: the matrix-core SpMM (mixed precision only; its
                                    32-bit offsets need n * padded batch * 8 bytes and
                                    64 bytes per union slot below 2^31, else
                                    DQ_ERR_INVALID_ARG) */
    int32_t spmm_traj_per_lane;  /* 1 or 2 (2: FP32 only, padded batch a multiple of 128) */
    int32_t spmm_phases;         /* trajectory slices of one SpMM run as 1, 2 or 4 launches */
    int32_t tile_rows;           /* 32, 64, 128 or 256 rows per workgroup of the update, combine
                                    and noise kernels */
    int32_t spmm_pipeline;       /* 0 or 1: gathers of the next 8 union columns in flight while
                                    the current 8 are multiplied in */
    int32_t concurrent_slices;   /* 1, 2 or 4: a graph-captured step runs that many equal slices
                                    of the trajectory slots on their own streams (each slice a
                                    multiple of 64 slots) */
    int32_t spmm_window_rows;    /* 0: the SpMMs above.  T > 0 (a multiple of 128, <= 4096): the ring
                                    SpMM (mixed precision only): a workgroup walks T rows for
                                    16 trajectories, holding the x rows they reference in an
                                    LDS ring of 1024 rows and multiplying 4-row blocks on the
                                    matrix cores; spmm_rows_per_wave and spmm_traj_per_lane
                                    then only shape the other kernels.  A Hamiltonian whose
                                    rows r .. r + 255 reference x rows more than 1024 apart
                                    (after rounding to groups of 4) is refused with
                                    DQ_ERR_INVALID_ARG, by dq_set_tuning or, with the ring
                                    chosen, by dq_system_set_hamiltonian */
    int32_t spmm_window_waves;   /* 0 (16), 4, 8 or 16 waves per workgroup of the ring SpMM */
    int32_t spmm_exact_sums;     /* matrix-core SpMMs (mixed precision; the row-blocked ones
                                    always sum in FP64): 0 = FP32 sums over 32 union slots
                                    moved into FP64 sums, 1 = FP64 sums after every 4 slots
                                    (about a third more SpMM time; over 5000 steps of the
                                    12,000-site tube 1.7e-7 instead of 1.6e-6 population
                                    error).  Ring SpMM: FP32 sums of 8 (0) or 4 (1) products
                                    per output moved into FP64 sums.  dq_autotune keeps the
                                    value it finds. */
} dq_tuning;

dq_status dq_get_tuning(dq_system *sys, dq_tuning *out);
/* Validates every field; drops a captured graph. */
dq_status dq_set_tuning(dq_system *sys, const dq_tuning *t);
/* Times every valid dq_tuning on the system's current Hamiltonian and batch
 * (reps cycles each: the noise prologue once the noise is set up, the
 * Lanczos kernels and the combine; psi, the noise offsets and the diagonal
 * are saved and restored, so the state propagated by dq_run is left as it
 * was), applies the fastest and returns it in *chosen.  Allocates a scratch
 * copy of that state for the duration of the call.  With cache_path, a
 * line keyed by device, n, nnz, padded batch, m and precision is looked up
 * first and, after a search, appended.  Needs set_hamiltonian. */
dq_status dq_autotune(dq_system *sys, int32_t reps, const char *cache_path, dq_tuning *chosen);

/* ---- files of the reference's workflow (host only) -------------------- */
/* Hamiltonian file (sparseMatrixClasses.cu:74-267): line 1 "%%MatrixMarket
 * matrix coordinate real general|symmetric|hermitian", line 2 BASE_ZERO or
 * BASE_ONE, line 3 "N N nnz", then nnz lines "row col value".  Symmetric and
 * hermitian files hold one triangle (the other is mirrored); a general file
 * must be symmetric.  Diagonal entries go to e_site[N], off-diagonal ones to a
 * column-sorted CSR (duplicates summed), ready for
 * dq_system_set_hamiltonian_csr.  Call with rowptr == NULL for *n_out and
 * *nnz_out (off-diagonal count) first. */
dq_status dq_read_hamiltonian_file(const char *path, int32_t *n_out, int64_t *nnz_out, int32_t *rowptr,
                                   int32_t *col, double *val, double *e_site);
/* Initial state file (systemClass.cu:1383-1481): BASE_ZERO|BASE_ONE, one
 * comment line, "index coefficient" lines up to END or the end of the file;
 * real coefficients (read as double), normalised; psi[2n] interleaved complex. */
dq_status dq_read_state_file(const char *path, int32_t n, double *psi);
/* Transition dipoles (spectraClass.cu:173-262): n lines "mu_x mu_y mu_z"; mu[3n]. */
dq_status dq_read_dipole_file(const char *path, int32_t n, double *mu);
/* Site positions for circular dichroism, the [posfile] of infile.in
 * (spectraClass.cu:296-360): n lines "x y z"; pos[3n]. */
dq_status dq_read_position_file(const char *path, int32_t n, double *pos);
/* Circular-dichroism vectors (spectraClass.cu:296-413): rd[i] = pos[i] x
 * mu[i] per site, mu_tot[d] = |mu_d| and rd_tot[d] = |rd_d| over the sites
 * for d = x, y, z.  Direction d of the CD run propagates psi(0) =
 * rd_d / rd_tot[d] and records C(n) against the bra mu_d / mu_tot[d]
 * (spectraClass.cu:463-496, 530-537).  Host code only. */
dq_status dq_cd_vectors(int32_t n, const double *pos, const double *mu, double *rd, double *mu_tot,
                        double *rd_tot);
/* Per-site parameter columns: ka_paramsfile "tau sigma" (ncol 2,
 * systemClass.cu:980-1071) or hsr_stddevfile "sigma" (ncol 1,
 * systemClass.cu:910-975); out[n][ncol]. */
dq_status dq_read_site_columns(const char *path, int32_t n, int32_t ncol, double *out);
/* Spectral density (systemClass.cu:1090-1160): "domega nosc", then nosc lines
 * "omega J" (cm^-1).  omega/J may be NULL to query *nosc_out. */
dq_status dq_read_spectral_density(const char *path, int32_t *nosc_out, double *domega_out, double *omega,
                                   double *J);

/* infile.in keys used by the propagation (systemClass.cu:183-735).  Lines
 * "key = value"; a key containing '!' is a comment.  (The reference drops the
 * character after '=', so it needs "key = value" with a blank; both forms
 * read the same here.) */
typedef struct dq_infile {
    double dt;                  /* dt (fs) */
    double temperature;         /* temperature (K), ZZ noise */
    double sigma_static;        /* sigmastatdis (cm^-1) */
    double gauss_width;         /* gausswidth (fs) */
    int32_t total_steps;        /* totalsteps */
    int32_t ensemble_runs;      /* ensembleruns */
    int32_t ensemble_mode;      /* ensembleavgmode: 0 NONE, 1 FIRST, 2 CONTINUATION */
    int32_t m;                  /* m */
    int32_t ka_substeps;        /* ka_substeps */
    int32_t noise_model;        /* simtype as a dq_noise value; CONST_HAM gives DQ_NOISE_NONE */
    int32_t prop_mode;          /* propmode: 0 PROP_MODE_LANCZOS, 1 PROP_MODE_TAYLOR_LANCZOS */
    int32_t static_dis;         /* staticdis */
    int32_t absorb_spec, cd_spec, lineshape_dress;
    int32_t state_looping, steps_per_loop;
    int32_t write_pops, print_all_pops, write_binary_pops, write_states, print_partic, debug_diags;
    int32_t print_pops[10];     /* printpop0..printpop9, -1 = unused */
    int32_t init_state, init_from_file;
    char ham_file[512], dipole_file[512], init_state_file[512], ka_params_file[512], hsr_file[512],
        zz_specd_file[512], out_folder[512];
    char pos_file[512];         /* posfile (circular dichroism) */
    char tk_mat_file[512];      /* tkmatfile: Taylor-Krylov auxiliary matrix */
    int32_t tk_p, tk_m_aux;     /* tk_p (default 10), tk_m_aux (default 24) */
} dq_infile;
dq_status dq_parse_infile(const char *path, dq_infile *out);

/* Population file (popPrintClass.cu:159-335, 442-522), text or binary.  Text:
 * "numsims S", "dt", "totalsteps T", "numprintedsites P", then for t = 1..T a
 * "STEP t" line and P lines "%25.17g (site)".  Binary: int32 S, double dt,
 * int32 T, int32 P, T*P doubles.  pops[t-1][site] are population sums over S
 * runs.  Writing to an existing file continues it (ENS_AVG_MODE_CONTINUATION):
 * its populations and run count are added after checking dt, T and P, and
 * the result replaces it through a temporary file. */
dq_status dq_write_pops_file(const char *path, int32_t binary, double dt, int32_t total_steps,
                             int32_t nsites, const double *pops, int32_t nsims);
/* pops may be NULL to read the header only */
dq_status dq_read_pops_file(const char *path, int32_t binary, int32_t *nsims, double *dt,
                            int32_t *total_steps, int32_t *nsites, double *pops);

#ifdef __cplusplus
}
#endif
#endif
