/*
 * dglm_oracle.h -- CPU restatement of the reference DGLM RHS + LSRK path.
 * TEST INFRASTRUCTURE ONLY (the checker; never part of the product library).
 */
#ifndef DGLM_ORACLE_H
#define DGLM_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* The datain/ array set of the reference (dglm_3d_launcher.py:74-149) with
 * operators row-major, plus the advection velocity per node. */
typedef struct oracle_problem {
  int N, Np, Nfp, K, nvar;
  double cfl, gamma, tau_scale, tmax;
  const int *vmapP, *vmapF, *Fmask;
  const double *WDr, *WDs, *WDt, *LIFT;
  const double *Fscale;
  const double *rx, *ry, *rz, *sx, *sy, *sz, *tx, *ty, *tz;
  const double *nx, *ny, *nz;
  const double *vx, *vy, *vz;
  /* Dirichlet faces (EtoB == 1, the reference's stubbed branch of
   * calc_fluxLM): ub = gD there, [nvar][K*4*Nfp]; NULL for none. */
  const int *EtoB;
  const double *gD;
  /* Numerical flux: 0 = the reference's LM flux (euler_dglm.cpp:110-183);
   * 1 = the textbook Rusanov (local Lax-Friedrichs) flux for Euler and the
   * exact upwind flux for advection, the configurations' own fluxes.  On a
   * Dirichlet face the prescribed state gD is the exterior trace u+. */
  int flux;
} oracle_problem;

/* One RHS at state u ([nvar][K*Np]); maxv, tau [K*4] optional. */
int oracle_rhs(const oracle_problem *P, const double *u, double *rhs, double *maxv, double *tau);

/* RHS at u with the face penalty tau [K*4] given instead of computed. */
int oracle_rhs_tau(const oracle_problem *P, const double *u, const double *tau, double *rhs);

/* Tangent-linear RHS with tau frozen: drhs = dR/du(u; tau) du (LM flux only;
 * -2 for flux != 0).  tau [K*4] is
 * the penalty to hold (the one of the step start inside a Runge-Kutta step);
 * NULL takes the tau that u itself gives. */
int oracle_rhs_tl(const oracle_problem *P, const double *u, const double *du, const double *tau, double *drhs);

/* nsteps of runRK on (u, k) in place; t and tstep advance; steps once t has
 * reached tmax are skipped (dt_hist entry 0). */
int oracle_run(const oracle_problem *P, double *u, double *k, int nsteps, double *t, int *tstep, double *dt_hist);

#ifdef __cplusplus
}
#endif

#endif
