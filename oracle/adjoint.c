/*
 * adjoint.c -- oracle reverse-mode gradient.  TEST INFRASTRUCTURE.
 * (filled in with the exact reverse sweep of krylov.c)
 */
#include "dq_oracle.h"
#include <math.h>

double orc_adjoint(const orc_params *p, const orc_site_params *sp,
                   const int32_t *rowptr, const int32_t *col, const double *val,
                   const double *psi0, uint64_t traj, int64_t nsteps,
                   const int32_t *group, int ngroups, const double *w_pop,
                   const double *bra, double w_spec, const double *corr_target,
                   int ckpt, double *grad_e, double *grad_g)
{
    (void)p; (void)sp; (void)rowptr; (void)col; (void)val; (void)psi0; (void)traj;
    (void)nsteps; (void)group; (void)ngroups; (void)w_pop; (void)bra; (void)w_spec;
    (void)corr_target; (void)ckpt; (void)grad_e; (void)grad_g;
    return NAN;
}
