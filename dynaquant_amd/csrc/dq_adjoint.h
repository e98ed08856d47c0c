// dq_adjoint.h -- tape and adjoint buffers of the reverse sweep (batched FP64).
#pragma once
#include "dq_internal.h"

namespace dq {

struct AdjDev {
    int m, K, last_c;
    size_t nB;                 // n * Bp
    size_t vs;                 // stride of the basis slots in the tape (the engine's Dev::vs)
    int Bp;
    // tape of one checkpoint segment, cycle c = (step - s0) * K + k
    double2* tapeV;            // [ncyc][m+1][vs] basis (slot 0 = cycle input psi)
    double* sc;                // [ncyc][m+1][Bp]
    double* alpha;             // [ncyc][m][Bp]
    double* beta;              // [ncyc][m][Bp]
    double* nrm2;              // [ncyc][Bp]   |psi'|^2 of the cycle
    double* tdiag;             // [ckpt][nB]   diagonal of each step
    double2* final_psi;        // [nB] normalised output of the segment's last cycle
    // adjoint state
    double2* pbar;             // adjoint of the cycle's normalised output
    double2* pbar1;            // adjoint of psi' (before renormalisation)
    double2* wb;               // wbar'
    double2* qb[3];            // ring of qbar_{j+1}, qbar_j, qbar_{j-1}
    double* grad_e;            // [nB] dL/d e_site
    double* gpart;             // [ntile_s][Bp] partial dL/dg
    double2* y;                // [m][Bp] Pade first column
    double* abar;              // [m][Bp]
    double* bbar;              // [m][Bp]
    double* r2;                // [Bp]
    double* r3;                // [m][Bp]
    double2* cbar;             // [T+1][Bp] adjoint of C(n)
    double2* pade_scratch;     // per trajectory matrices
    size_t pade_stride;
    int max_s;

    __host__ __device__ const double2* basis(int c) const { return tapeV + (size_t)c * (m + 1) * vs; }
    __host__ __device__ const double2* psi_out(int c) const { return c < last_c ? basis(c + 1) : final_psi; }
    __host__ __device__ double2* qbar_slot(int k) const { return qb[k % 3]; }
    __host__ __device__ const double* diag_step(int c) const { return tdiag + (size_t)(c / K) * nB; }
};

hipError_t adj_renorm(const Dev& d, const AdjDev& a, int c, hipStream_t s);
hipError_t adj_pade(const Dev& d, const AdjDev& a, int c, hipStream_t s);
hipError_t adj_lanczos(const Dev& d, const AdjDev& a, int c, hipStream_t s);
hipError_t adj_add_corr(const Dev& d, const AdjDev& a, int64_t step, double2* dst, hipStream_t s);

}  // namespace dq
