/*
 * tube.c -- oracle restatement of the Hamiltonian setup.  TEST INFRASTRUCTURE.
 *
 * The reference reads a precomputed COO Hamiltonian (sparseMatrixClasses.cu:
 * 74-267) and converts it to CSR (:489-540); the example Hamiltonian file is
 * not shipped.  The benchmark systems are cylindrical aggregates of point
 * dipoles, so the oracle builds them by brute force over all site pairs,
 * independently of the engine's windowed generator.
 */
#include "dq_oracle.h"
#include <math.h>
#include <stdlib.h>

int orc_tube_sites(int n_walls, const int *rings, const double *radius,
                   int per_ring, double ring_dz, double mu_debye,
                   double tilt_deg, int order,
                   double *pos, double *dip, int *wall_of_site)
{
    const double pi = atan(1.0) * 4.0;
    int n = 0;
    for (int w = 0; w < n_walls; ++w) n += rings[w] * per_ring;
    if (order == 1)
        for (int w = 1; w < n_walls; ++w)
            if (rings[w] != rings[0]) return -1;
    double tilt = tilt_deg * pi / 180.0;
    double ct = cos(tilt), st = sin(tilt);
    int wall_off = 0;
    for (int w = 0; w < n_walls; ++w) {
        for (int r = 0; r < rings[w]; ++r) {
            for (int k = 0; k < per_ring; ++k) {
                int i;
                if (order == 0) i = wall_off + r * per_ring + k;
                else i = r * (n_walls * per_ring) + w * per_ring + k;
                double phi = 2.0 * pi * (double)k / (double)per_ring;
                if (r & 1) phi += pi / (double)per_ring;
                double cp = cos(phi), sp = sin(phi);
                if (pos) {
                    pos[3 * i + 0] = radius[w] * cp;
                    pos[3 * i + 1] = radius[w] * sp;
                    pos[3 * i + 2] = ring_dz * (double)r;
                }
                if (dip) {
                    dip[3 * i + 0] = mu_debye * (ct * -sp);
                    dip[3 * i + 1] = mu_debye * (ct * cp);
                    dip[3 * i + 2] = mu_debye * st;
                }
                if (wall_of_site) wall_of_site[i] = w;
            }
        }
        wall_off += rings[w] * per_ring;
    }
    return n;
}

static double pair_j(const double *pi_, const double *pj, const double *mi,
                     const double *mj, double kappa, double *r_out)
{
    double dx = pj[0] - pi_[0], dy = pj[1] - pi_[1], dz = pj[2] - pi_[2];
    double r2 = dx * dx + dy * dy + dz * dz;
    double r = sqrt(r2);
    *r_out = r;
    double ux = dx / r, uy = dy / r, uz = dz / r;
    double mm = mi[0] * mj[0] + mi[1] * mj[1] + mi[2] * mj[2];
    double mir = mi[0] * ux + mi[1] * uy + mi[2] * uz;
    double mjr = mj[0] * ux + mj[1] * uy + mj[2] * uz;
    return kappa * (mm - 3.0 * mir * mjr) / (r2 * r);
}

int64_t orc_couplings_csr(int n, const double *pos, const double *dip,
                          double cutoff_nm, double kappa,
                          int32_t *rowptr, int32_t *col, double *val)
{
    int64_t nnz = 0;
    if (rowptr) rowptr[0] = 0;
    for (int i = 0; i < n; ++i) {
        for (int j = 0; j < n; ++j) {
            if (j == i) continue;
            double r;
            double J = pair_j(pos + 3 * i, pos + 3 * j, dip + 3 * i, dip + 3 * j, kappa, &r);
            if (!(r < cutoff_nm)) continue;
            if (val) { col[nnz] = j; val[nnz] = J; }
            ++nnz;
        }
        if (rowptr) rowptr[i + 1] = (int32_t)nnz;
    }
    return nnz;
}
