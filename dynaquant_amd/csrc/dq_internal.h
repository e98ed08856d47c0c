// dq_internal.h -- device state shared by the host driver (dq_api.cpp) and
// the kernels (dq_kernels.hip).  Layouts in HBM:
//   per-(site, trajectory) arrays are [i][b] with b contiguous (Bp = padded B),
//   so one CSR nonzero gathers a contiguous run of trajectories;
//   Krylov basis slot k (k = 0..m) is one such [n][Bp] array;
//   per-trajectory scalars are [k][Bp];
//   partial sums are [q][tile][Bp].
#pragma once
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace dq {

constexpr int kMaxM = 32;          // Krylov depth limit (Pade matrix <= 34)
constexpr int kMaxGroups = 8;      // population groups reduced on the device
constexpr int kThreads = 256;
constexpr int kRBMin = 2;          // rows per wavefront of the row-blocked SpMM: 2, 4 or 8
constexpr int kRowsMin = 32;       // smallest elementwise tile (rows per workgroup)
constexpr int kRingCap = 1024;     // ring SpMM: x rows held per column in LDS (a power of 2)
// Bytes of the newest basis vectors the batched update keeps cached: 7/8 of
// MI355X's 256 MB last-level (memory-side) cache.  Those vectors are the
// ones the next iterations re-read soonest; the older ones are streamed with
// the nontemporal hint so that they do not evict them.
constexpr size_t kUpdCacheBytes = (size_t)224 << 20;
constexpr int kRingChunk = 128;    // ring SpMM: rows multiplied between two ring refills
constexpr int kRingMaxRows = 4096; // ring SpMM: most rows per workgroup

struct Dev {
    int n, Bp, B, m;
    int mixed;                      // 1: float2 basis and FP32 SpMM
    int plain;                      // 1: B == 1 single-trajectory kernels
    int rt_e, ntile_e;              // elementwise kernels: rows per tile, tiles
    int rt_o, ntile_o;              // prologue (noise, observables) and its reduction: rows
                                    // per tile, tiles (kRowsMin-row tiles in the batched path:
                                    // the noise recursion is latency-bound and needs the
                                    // workgroups a concurrent slice of 64 trajectories lacks
                                    // with longer tiles)
    int rt_s, ntile_s;              // SpMM kernels
    int ntile_p;                    // tiles of the SpMM partials in part_s (ntile_s, or the
                                    // ring kernel's row ranges when ring_rows > 0)
    int rb;                         // batched SpMM: rows per wavefront (row block)
    int spmm_tpl;                   // batched FP32 SpMM: trajectories per lane (1 or 2)
    int spmm_phases;                // batched SpMM: trajectory slices run as this many launches
    int spmm_pipe;                  // batched SpMM: 1 = two gather buffers in flight
    int spmm_exact;                 // matrix-core SpMM: 1 = FP64 sums after every K-step
    int s0, sn;                     // batched kernels work on trajectory slots [s0, s0 + sn)
    int upd_cached;                 // batched update: the newest upd_cached groups of 4 basis
                                    // vectors are read with cached loads, older ones with
                                    // nontemporal loads (kUpdCacheBytes over the active sites)
    int S;                          // OU substeps
    int noise_model;                // DQ_NOISE_*: 0 OU, 1 HSR, 2 ZZ real, 3 none
    int zz_nosc;                    // ZZ oscillators per site
    double dt;                      // fs (ZZ time t = step dt)
    const double* hsr_sigma;        // [n] HSR standard deviations
    const double* zz_w2pi;          // [zz_nosc] omega / 2 pi, 1/fs
    const double* zz_coeff;         // [zz_nosc] cm^-1
    uint64_t seed;
    double subdt, sqrt_subdt;
    double g;                       // coupling scale
    double theta;                   // tau / hbar  (A = -i theta Tbar)
    double tol;                     // breakdown tolerance
    // Hamiltonian
    const int32_t* rowptr;
    const int32_t* col;
    const double* val64;
    const float* val32;
    // row-blocked copy for the batched SpMM: blocks of rb rows, the union
    // of their columns and rb values per union column (0 where absent)
    const int32_t* rb_ptr;          // [n/rb + 1] offsets into rb_col
    const int32_t* rb_col;          // [total union columns]
    const float* rb_val32;          // [total][rb]
    const double* rb_val64;         // [total][rb]
    const uint16_t* rb_lo16;        // rb = 16: bfloat16 low parts of rb_val32 (16-slot MFMA operands)
    // ring SpMM (dq_ring.hip), mixed precision: a workgroup walks ring_rows
    // rows (a multiple of kRingChunk; 0 = off) for 16 trajectories with
    // ring_w waves; lists built from the CSR for 4-row blocks
    int ring_rows, ring_w;
    const int32_t* ring_item;       // [ceil(n/4) + 1] first item (up to 16 column pairs) of each 4-row block
    const float* ring_a;            // [item][64][2] FP32 couplings: lane 32h + 4a + i, row i, column
                                    // 2 p + h of pair a (.x) and pair 8 + a (.y) of the item
    const uint2* ring_alo;          // [item][group 4][h][row 4] bfloat16 low parts of the 4 pairs
    const uint4* ring_off;          // [item][4] LDS byte offsets (2 p mod kRingCap) * 4 of the 16 pairs;
                                    // bits 0-2 of the first: pairs / 4 - 1, last item of the block
    const int2* ring_win;           // [ceil(n/kRingChunk)] window rows [lo, hi) of each 128-row chunk
    const double* inv_tau;          // [n]
    const double* coeff;            // [n] sigma*sqrt(2/tau)
    const double* ou_sigma;         // [n]
    const double* e_site;           // [n]
    double sigma_static;
    // trajectories
    const int32_t* nact;            // [Bp]
    const uint64_t* traj_id;        // [Bp]
    // state
    double2* psi;                   // [n][Bp]
    void* V;                        // (m+1) slots of [n][Bp] float2 / double2, vs elements apart
    size_t vs;                      // slot stride: n Bp rounded up to 64, + 32, so that the
                                    // slots are an odd multiple of 256 B apart (a multiple of
                                    // 16 KB puts every slot's row i on the same HBM channels)
    void* diag;                     // [n][Bp] float / double
    double* de;                     // [n][Bp] OU offsets
    double* e0;                     // [n][Bp] static site energies
    // scalars
    double* sc;                     // [m+1][Bp] scale of basis slot k
    double* alpha;                  // [m][Bp]
    double* beta;                   // [m][Bp]
    double2* hco;                   // [m][m][Bp] projection coefficients h^(j)_k
    double2* gram;                  // [m+1][Bp] Gram row of the newest slot
    double* nrm2;                   // [Bp] |psi|^2 after combine
    double2* ycoef;                 // [m][Bp] y_k * s_k
    double* err_sum;                // [Bp]
    double* err_max;                // [Bp] largest clamped estimate of a cycle (substep control), or null
    int* broke;                     // [Bp]
    int* status;                    // [1] sticky error flags
    double* part;                   // partial sums
    double* part_s;                 // batched SpMM partials <v_j|w> [q < 2][tile][Bp]
    double* red;                    // [q][Bp] reduced sums
    // observables
    const int32_t* group;           // [n]
    int ngroups, site_mode, every;
    int64_t max_steps;
    double* pops;                   // [rec][g][Bp] or [rec][i][Bp] (site mode)
    const double2* bra;             // [nbra][n] or null
    const int32_t* bra_idx;         // [Bp] row of the bra table of each slot, or null (row 0)
    double2* corr;                  // [max_steps+1][Bp]
    double* partic;                 // [max_steps+1][Bp] participation ratio, or null
    // Taylor-Krylov propagation (taylorKrylovClass.cu:133-434)
    int aug;                        // augmentation of Tbar: 2, or p for the auxiliary Krylov part
    int tk;                         // 1: this Dev is the auxiliary Krylov part of a Taylor-Krylov cycle
    int tk_p;                       // Taylor order p (0: Lanczos propagation)
    double2* tk_t;                  // [n][Bp] Taylor part sum_{k<p} c^k A^k psi / k!
    double2* tk_u;                  // 2 x [n][Bp] A^k psi, alternating
    const double2* tk_coef;         // [p] c^k / k!
    double* tk_nrm;                 // [Bp] |A^p psi|
    double* tk_err;                 // [Bp] Expokit estimate of the auxiliary part (unclamped)
    int64_t* step;                  // device step counter
};

// any observable recorded in the prologue
__host__ __device__ inline bool obs_on(const Dev& d)
{
    return d.every > 0 || d.bra != nullptr || d.partic != nullptr;
}

enum StatusBits { kStatusScaling = 1, kStatusNonfinite = 2 };

// first row of red[] holding the reduced <v_j|w> when a reduction launch sums
// it (after the Gram sums)
__host__ __device__ inline int kRedC(int m) { return 2 * (m + 1) + 1; }

// Batched Lanczos iteration j: the matrix-core SpMM carries the reduction of
// update j-1's Gram row (2j+1 quantities per group of 64 slots) in leading
// blocks of its own launch, padded to a multiple of 8 so that its row blocks
// keep their XCDs (and the scalars kernel sums the SpMM tiles); the other
// SpMMs leave both to launch_reduce_all.
__host__ __device__ inline bool spmm_fuses_gram(const Dev& d) { return !d.plain && d.rb == 16 && d.ring_rows <= 0; }
__host__ __device__ inline int gram_task_blocks(const Dev& d, int j, int last)
{
    if (last || j < 1 || !spmm_fuses_gram(d)) return 0;
    return ((d.sn / 64) * (2 * j + 1) + 7) & ~7;
}

// launch wrappers (dq_kernels.hip)
hipError_t launch_init_disorder(const Dev& d, hipStream_t s);
hipError_t launch_prologue(const Dev& d, int renorm, int noise, int record, int copy_v0, hipStream_t s);
hipError_t launch_spmm(const Dev& d, int j, int last, hipStream_t s);
hipError_t launch_scalars_h(const Dev& d, int j, hipStream_t s);
hipError_t launch_update(const Dev& d, int j, hipStream_t s);
hipError_t launch_reduce(const Dev& d, int nq, int ntile, hipStream_t s);
// batched variants (dq_batched.hip)
hipError_t launch_spmm_b(const Dev& d, int j, int last, hipStream_t s);
hipError_t launch_scalars_hb(const Dev& d, int j, hipStream_t s);
hipError_t launch_reduce_wide(const Dev& d, int nq, int ntile, int dst, hipStream_t s);
// Lanczos iteration j, batched, when the SpMM does not carry the Gram
// reduction (spmm_fuses_gram): <v_j|w> over the SpMM tiles (part_s) into
// red[kRedC(m)..+1] and, for j >= 1, the Gram row and |v_j|^2 of update j-1
// (part, 2j+1 quantities) into red[0..2j], in one launch
hipError_t launch_reduce_all(const Dev& d, int j, hipStream_t s);
hipError_t launch_update_b(const Dev& d, int j, hipStream_t s);

// Timing events of the launch in progress (per-kernel profiling): a launcher
// that takes them launches through hipExtLaunchKernelGGL, whose events carry
// the kernel's own start and end (no dispatch gap), and sets `used`.
struct KernelEvents {
    hipEvent_t start = nullptr, stop = nullptr;
    bool used = false;
};
extern thread_local KernelEvents g_kev;

// launch `kernel` with the profiling events when there are any; false: the
// caller launches it the ordinary way
template <typename F, typename... Args>
inline bool launch_with_events(F kernel, dim3 grid, dim3 block, hipStream_t s, Args... args)
{
    if (!g_kev.start) return false;
    hipExtLaunchKernelGGL(kernel, grid, block, 0, s, g_kev.start, g_kev.stop, 0, args...);
    g_kev.used = true;
    return true;
}
// ring SpMM (dq_ring.hip)
hipError_t launch_spmm_ring(const Dev& d, int j, int last, hipStream_t s);
size_t spmm_ring_lds_bytes(int waves, int rows);
hipError_t spmm_ring_set_lds_limit();
hipError_t launch_combine_b(const Dev& d, hipStream_t s);
hipError_t launch_pade(const Dev& d, hipStream_t s);
hipError_t launch_combine(const Dev& d, hipStream_t s);
hipError_t launch_finish_norm(const Dev& d, int advance_step, hipStream_t s);
hipError_t launch_reduce_obs(const Dev& d, hipStream_t s);
hipError_t launch_set_unit_norm(const Dev& d, hipStream_t s);
hipError_t launch_advance_step(const Dev& d, hipStream_t s);
hipError_t launch_pade_tbar(int m, int count, const double* alpha, const double* beta,
                            double theta, double2* E, int* s_out, hipStream_t s);
hipError_t launch_fft(int n, int batch, double2* a, double2* b, int sign, hipStream_t s);
// Taylor-Krylov (dq_taylor.hip)
hipError_t launch_tk_power(const Dev& d, int k, hipStream_t s);
hipError_t launch_tk_power_b(const Dev& d, int k, hipStream_t s);
hipError_t launch_tk_start(const Dev& d, hipStream_t s);
hipError_t launch_tk_combine(const Dev& d, hipStream_t s);
hipError_t launch_spectrum_prep(const Dev& d, int64_t T, double dt, double sigma, int nfft,
                                double2* buf, const double* wts, hipStream_t s);
// dq_noalloc.hip: the current device's fixed workspace of 2 * kNoAllocMaxNfft
// complex doubles (its address, looked up once per device)
constexpr int kNoAllocMaxNfft = 1 << 17;
hipError_t noalloc_scratch(double2** out);

}  // namespace dq
