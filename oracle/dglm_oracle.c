/*
 * dglm_oracle.c -- CPU restatement of the reference explicit DGLM path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker for the HIP product
 * path: it is built into oracle/liboracle.so and loaded only by tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg.  The product
 * library (libexadglm_hip.so) never links or calls it.
 *
 * It restates, with runtime N / NVAR and OpenMP, the single-partition CPU
 * solver of the reference:
 *   calc_maxv    dim3/src/cpp/advec_dglm.cpp:6-35,  euler_dglm.cpp:6-50
 *   calc_dt      advec_dglm.cpp:38-53,              euler_dglm.cpp:53-69
 *   update_tau   advec_dglm.cpp:56-66,              euler_dglm.cpp:72-84
 *   update_ub    advec_dglm.cpp:69-85,              euler_dglm.cpp:87-107
 *   calc_fluxLM  advec_dglm.cpp:88-133,             euler_dglm.cpp:110-183
 *   calc_fluxes  advec_dglm.cpp:136-169,            euler_dglm.cpp:186-251
 *   update_rhs   advec_dglm.cpp:172-195,            euler_dglm.cpp:254-302
 *   lserk_stage  advec_time.cpp:15-26,              euler_time.cpp:15-38
 *   runRK        advec_time.cpp:29-97,              euler_time.cpp:41-107
 * with the same floating-point expression order, so that it agrees with the
 * compiled reference to rounding (pinned in tests/test_oracle.py).  Faces
 * with EtoB == 1 take a prescribed ub (the reference's Dirichlet stub).
 *
 * Beside the reference's LM flux it has the configurations' own surface
 * fluxes (flux = 1: Rusanov for Euler, exact upwind for advection), written
 * from the textbook formulas (calc_flux_rusanov) and pinned by a numpy
 * restatement of those formulas and by their equality with the LM flux on
 * continuous traces (tests/test_oracle.py).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "dglm_oracle.h"

static const double RKA[5] = {0.0, -567301805773.0 / 1357537059087.0, -2404267990393.0 / 2016746695238.0,
                              -3550918686646.0 / 2091501179385.0, -1275806237668.0 / 842570457699.0};
static const double RKB[5] = {1432997174477.0 / 9575080441755.0, 5161836677717.0 / 13612068292357.0,
                              1720146321549.0 / 2090206949498.0, 3134564353537.0 / 4481467310338.0,
                              2277821191437.0 / 14882151754819.0};

#define NFACE 4

static void calc_maxv(const oracle_problem *P, const double *u, double *maxv, double *maxv_fs) {
  const int K = P->K, Np = P->Np, Nfp = P->Nfp;
  const size_t KNp = (size_t)K * Np;
  const double G = P->gamma;
#pragma omp parallel for
  for (int fidx = 0; fidx < K * NFACE; fidx++) {
    int ei = fidx / NFACE, fi = fidx % NFACE;
    double max_v2 = 0, max_c2 = 0;
    for (int j = 0; j < Nfp; j++) {
      size_t idxM = (size_t)ei * Np + P->Fmask[fi * Nfp + j];
      if (P->nvar == 5) {
        double u1 = u[idxM], u2 = u[KNp + idxM], u3 = u[2 * KNp + idxM], u4 = u[3 * KNp + idxM],
               u5 = u[4 * KNp + idxM];
        double vx = u2 / u1, vy = u3 / u1, vz = u4 / u1;
        double v2 = vx * vx + vy * vy + vz * vz;
        if (v2 > max_v2) max_v2 = v2;
        double pres = (G - 1) * (u5 - 0.5 * u1 * v2);
        double sound_c2 = fabs(G * pres / u1);
        if (sound_c2 > max_c2) max_c2 = sound_c2;
      } else {
        double vx = P->vx[idxM], vy = P->vy[idxM], vz = P->vz[idxM];
        double v2 = vx * vx + vy * vy + vz * vz;
        if (v2 > max_v2) max_v2 = v2;
      }
    }
    double mv = P->nvar == 5 ? sqrt(max_v2) + sqrt(max_c2) : sqrt(max_v2);
    maxv[fidx] = mv;
    maxv_fs[fidx] = mv * P->Fscale[fidx];
  }
}

static double calc_dt(const oracle_problem *P, const double *maxv_fs) {
  double max_v = 0;
#pragma omp parallel for reduction(max : max_v)
  for (int fidx = 0; fidx < P->K * NFACE; fidx++) {
    double v = maxv_fs[fidx];
    if (v > max_v) max_v = v;
  }
  return (P->cfl / (2 * P->N + 1)) * (1 / max_v);
}

static void update_tau(const oracle_problem *P, const double *maxv, double *tau) {
#pragma omp parallel for
  for (int fidx = 0; fidx < P->K * NFACE; fidx++)
    tau[fidx] = 0.25 * (maxv[fidx] + maxv[P->vmapF[fidx]]) * P->tau_scale;
}

static void update_ub(const oracle_problem *P, const double *u, double *ub) {
  const int K = P->K, Np = P->Np, Nfp = P->Nfp, NFF = NFACE * Nfp;
  const size_t KNp = (size_t)K * Np, KNF = (size_t)K * NFF;
#pragma omp parallel for
  for (long idx = 0; idx < (long)K * NFF; idx++) {
    int ei = (int)(idx / NFF), fi = (int)((idx / Nfp) % NFACE);
    size_t idxM = (size_t)ei * Np + P->Fmask[fi * Nfp + idx % Nfp];
    size_t idxP = (size_t)P->vmapP[idx];
    if (P->gD && P->EtoB && P->EtoB[idx / Nfp] == 1) {
      /* Dirichlet: the prescribed face state (euler_dglm.cpp:133-139,
         advec_dglm.cpp:113-119 leave it as "ub = function of x,y,z,t") */
      for (int v = 0; v < P->nvar; v++) ub[v * KNF + idx] = P->gD[v * KNF + idx];
      continue;
    }
    for (int v = 0; v < P->nvar; v++) ub[v * KNF + idx] = 0.5 * (u[v * KNp + idxM] + u[v * KNp + idxP]);
  }
}

static void calc_fluxLM(const oracle_problem *P, const double *tau, const double *ub, const double *u,
                        double *fluxLM) {
  const int K = P->K, Np = P->Np, Nfp = P->Nfp, NFF = NFACE * Nfp;
  const size_t KNp = (size_t)K * Np, KNF = (size_t)K * NFF;
  const double G = P->gamma;
#pragma omp parallel for
  for (int fidx = 0; fidx < K * NFACE; fidx++) {
    double nx = P->nx[fidx], ny = P->ny[fidx], nz = P->nz[fidx], fs = P->Fscale[fidx];
    int ei = fidx / NFACE, fi = fidx % NFACE;
    for (int j = 0; j < Nfp; j++) {
      size_t idx = (size_t)fidx * Nfp + j;
      size_t idxM = (size_t)ei * Np + P->Fmask[fi * Nfp + j];
      if (P->nvar == 5) {
        double ub1 = ub[idx], ub2 = ub[KNF + idx], ub3 = ub[2 * KNF + idx], ub4 = ub[3 * KNF + idx],
               ub5 = ub[4 * KNF + idx];
        double vb2 = ub2 * ub2 + ub3 * ub3 + ub4 * ub4;
        double F[5], Gf[5], W[5];
        F[0] = ub2;
        F[1] = (ub2 * ub2) / ub1 + (G - 1) * (ub5 - 0.5 * vb2 / ub1);
        F[2] = (ub2 * ub3) / ub1;
        F[3] = (ub2 * ub4) / ub1;
        F[4] = (ub5 + (G - 1) * (ub5 - 0.5 * vb2 / ub1)) * (ub2 / ub1);
        Gf[0] = ub3;
        Gf[1] = ub2 * ub3 / ub1;
        Gf[2] = (ub3 * ub3) / ub1 + (G - 1) * (ub5 - 0.5 * vb2 / ub1);
        Gf[3] = ub3 * ub4 / ub1;
        Gf[4] = (ub5 + (G - 1) * (ub5 - 0.5 * vb2 / ub1)) * (ub3 / ub1);
        W[0] = ub4;
        W[1] = ub2 * ub4 / ub1;
        W[2] = ub3 * ub4 / ub1;
        W[3] = (ub4 * ub4) / ub1 + (G - 1) * (ub5 - 0.5 * vb2 / ub1);
        W[4] = (ub5 + (G - 1) * (ub5 - 0.5 * vb2 / ub1)) * (ub4 / ub1);
        double ubv[5] = {ub1, ub2, ub3, ub4, ub5};
        for (int v = 0; v < 5; v++) {
          double nf = nx * F[v] + ny * Gf[v] + nz * W[v];
          fluxLM[v * KNF + idx] = (nf - tau[fidx] * (ubv[v] - u[v * KNp + idxM])) * fs;
        }
      } else {
        double b = ub[idx];
        double fF = P->vx[idxM] * b, fG = P->vy[idxM] * b, fW = P->vz[idxM] * b;
        double nf = nx * fF + ny * fG + nz * fW;
        fluxLM[idx] = (nf - tau[fidx] * (b - u[idxM])) * fs;
      }
    }
  }
}

/* Textbook numerical fluxes of the configurations (no reference counterpart;
 * the reference has only the LM flux).  Euler: Rusanov / local Lax-Friedrichs
 *   F* = (F(u-).n + F(u+).n)/2 + (lambda/2)(u- - u+),
 *   lambda = max over the two traces of |v.n| + sqrt(gamma p / rho)
 * (Toro, "Riemann Solvers and Numerical Methods for Fluid Dynamics", 3rd ed.,
 * sec. 10.5.1); advection: exact upwinding F* = (v.n) u-  if v.n > 0, else
 * (v.n) u+ (v taken at the interior node, as the LM flux does).  The result is
 * scaled by Fscale like fluxLM.  u+ is the neighbour's node value, or the
 * prescribed state gD on a Dirichlet face. */
static void euler_fn(const double U[5], double nx, double ny, double nz, double G, double Fn[5], double *lam) {
  const double rho = U[0], vx = U[1] / rho, vy = U[2] / rho, vz = U[3] / rho;
  const double p = (G - 1) * (U[4] - 0.5 * rho * (vx * vx + vy * vy + vz * vz));
  /* physical fluxes in x, y, z */
  const double Fx[5] = {rho * vx, rho * vx * vx + p, rho * vy * vx, rho * vz * vx, (U[4] + p) * vx};
  const double Fy[5] = {rho * vy, rho * vx * vy, rho * vy * vy + p, rho * vz * vy, (U[4] + p) * vy};
  const double Fz[5] = {rho * vz, rho * vx * vz, rho * vy * vz, rho * vz * vz + p, (U[4] + p) * vz};
  for (int v = 0; v < 5; v++) Fn[v] = nx * Fx[v] + ny * Fy[v] + nz * Fz[v];
  const double vn = nx * vx + ny * vy + nz * vz;
  *lam = fabs(vn) + sqrt(fabs(G * p / rho));
}

static void calc_flux_rusanov(const oracle_problem *P, const double *u, double *fluxLM) {
  const int K = P->K, Np = P->Np, Nfp = P->Nfp, NFF = NFACE * Nfp;
  const size_t KNp = (size_t)K * Np, KNF = (size_t)K * NFF;
  const double G = P->gamma;
#pragma omp parallel for
  for (int fidx = 0; fidx < K * NFACE; fidx++) {
    const double nx = P->nx[fidx], ny = P->ny[fidx], nz = P->nz[fidx], fs = P->Fscale[fidx];
    const int ei = fidx / NFACE, fi = fidx % NFACE;
    const int dir = P->gD && P->EtoB && P->EtoB[fidx] == 1;
    for (int j = 0; j < Nfp; j++) {
      const size_t idx = (size_t)fidx * Nfp + j;
      const size_t iM = (size_t)ei * Np + P->Fmask[fi * Nfp + j], iP = (size_t)P->vmapP[idx];
      if (P->nvar == 5) {
        double uM[5], uP[5], fm[5], fp[5], lm, lp;
        for (int v = 0; v < 5; v++) {
          uM[v] = u[v * KNp + iM];
          uP[v] = dir ? P->gD[v * KNF + idx] : u[v * KNp + iP];
        }
        euler_fn(uM, nx, ny, nz, G, fm, &lm);
        euler_fn(uP, nx, ny, nz, G, fp, &lp);
        const double lam = lm > lp ? lm : lp;
        for (int v = 0; v < 5; v++) fluxLM[v * KNF + idx] = (0.5 * (fm[v] + fp[v]) + 0.5 * lam * (uM[v] - uP[v])) * fs;
      } else {
        const double uM = u[iM], uP = dir ? P->gD[idx] : u[iP];
        const double vn = nx * P->vx[iM] + ny * P->vy[iM] + nz * P->vz[iM];
        fluxLM[idx] = (vn > 0 ? vn * uM : vn * uP) * fs;
      }
    }
  }
}

static void calc_fluxes(const oracle_problem *P, const double *u, double *fR, double *fS, double *fT) {
  const int K = P->K, Np = P->Np;
  const size_t KNp = (size_t)K * Np;
  const double G = P->gamma;
#pragma omp parallel for
  for (int ei = 0; ei < K; ei++) {
    double rx = P->rx[ei], ry = P->ry[ei], rz = P->rz[ei];
    double sx = P->sx[ei], sy = P->sy[ei], sz = P->sz[ei];
    double tx = P->tx[ei], ty = P->ty[ei], tz = P->tz[ei];
    for (int i = 0; i < Np; i++) {
      size_t idx = (size_t)ei * Np + i;
      if (P->nvar == 5) {
        double u1 = u[idx], u2 = u[KNp + idx], u3 = u[2 * KNp + idx], u4 = u[3 * KNp + idx], u5 = u[4 * KNp + idx];
        double v2 = u2 * u2 + u3 * u3 + u4 * u4;
        double F[5], Gf[5], W[5];
        F[0] = u2;
        F[1] = (u2 * u2) / u1 + (G - 1) * (u5 - 0.5 * v2 / u1);
        F[2] = (u2 * u3) / u1;
        F[3] = (u2 * u4) / u1;
        F[4] = (u5 + (G - 1) * (u5 - 0.5 * v2 / u1)) * (u2 / u1);
        Gf[0] = u3;
        Gf[1] = u2 * u3 / u1;
        Gf[2] = (u3 * u3) / u1 + (G - 1) * (u5 - 0.5 * v2 / u1);
        Gf[3] = u3 * u4 / u1;
        Gf[4] = (u5 + (G - 1) * (u5 - 0.5 * v2 / u1)) * (u3 / u1);
        W[0] = u4;
        W[1] = u2 * u4 / u1;
        W[2] = u3 * u4 / u1;
        W[3] = (u4 * u4) / u1 + (G - 1) * (u5 - 0.5 * v2 / u1);
        W[4] = (u5 + (G - 1) * (u5 - 0.5 * v2 / u1)) * (u4 / u1);
        for (int v = 0; v < 5; v++) {
          fR[v * KNp + idx] = rx * F[v] + ry * Gf[v] + rz * W[v];
          fS[v * KNp + idx] = sx * F[v] + sy * Gf[v] + sz * W[v];
          fT[v * KNp + idx] = tx * F[v] + ty * Gf[v] + tz * W[v];
        }
      } else {
        double fF = P->vx[idx] * u[idx], fG = P->vy[idx] * u[idx], fW = P->vz[idx] * u[idx];
        fR[idx] = rx * fF + ry * fG + rz * fW;
        fS[idx] = sx * fF + sy * fG + sz * fW;
        fT[idx] = tx * fF + ty * fG + tz * fW;
      }
    }
  }
}

static void update_rhs(const oracle_problem *P, const double *fR, const double *fS, const double *fT,
                       const double *fLM, double *rhs) {
  const int K = P->K, Np = P->Np, NFF = NFACE * P->Nfp;
  const size_t KNp = (size_t)K * Np, KNF = (size_t)K * NFF;
  /* Operators transposed once, so that the loop over output rows i is
   * contiguous and vectorises; every output keeps the reference's summation
   * order over j (acc += WDr r + WDs s + WDt t, then acc -= LIFT l), so the
   * values are those of the row-by-row loops of the reference. */
  double *T = malloc(((size_t)3 * Np * Np + (size_t)NFF * Np) * sizeof(double));
  if (!T) abort();
  double *Tr = T, *Ts = T + (size_t)Np * Np, *Tt = T + (size_t)2 * Np * Np, *Tl = T + (size_t)3 * Np * Np;
  for (int i = 0; i < Np; i++) {
    for (int j = 0; j < Np; j++) {
      Tr[(size_t)j * Np + i] = P->WDr[i * Np + j];
      Ts[(size_t)j * Np + i] = P->WDs[i * Np + j];
      Tt[(size_t)j * Np + i] = P->WDt[i * Np + j];
    }
    for (int j = 0; j < NFF; j++) Tl[(size_t)j * Np + i] = P->LIFT[i * NFF + j];
  }
#pragma omp parallel
  {
    double *acc = malloc((size_t)Np * sizeof(double));
    if (!acc) abort();
#pragma omp for
    for (int ei = 0; ei < K; ei++) {
      for (int v = 0; v < P->nvar; v++) {
        const double *r = fR + v * KNp + (size_t)ei * Np, *s = fS + v * KNp + (size_t)ei * Np,
                     *t = fT + v * KNp + (size_t)ei * Np, *l = fLM + v * KNF + (size_t)ei * NFF;
        for (int i = 0; i < Np; i++) acc[i] = 0;
        for (int j = 0; j < Np; j++) {
          const double rj = r[j], sj = s[j], tj = t[j];
          const double *a = Tr + (size_t)j * Np, *b = Ts + (size_t)j * Np, *c = Tt + (size_t)j * Np;
          for (int i = 0; i < Np; i++) acc[i] += a[i] * rj + b[i] * sj + c[i] * tj;
        }
        for (int j = 0; j < NFF; j++) {
          const double lj = l[j];
          const double *a = Tl + (size_t)j * Np;
          for (int i = 0; i < Np; i++) acc[i] -= a[i] * lj;
        }
        double *o = rhs + v * KNp + (size_t)ei * Np;
        for (int i = 0; i < Np; i++) o[i] = acc[i];
      }
    }
    free(acc);
  }
  free(T);
}

typedef struct {
  double *maxv, *maxv_fs, *tau, *ub, *fR, *fS, *fT, *fLM;
} work_t;

static int work_alloc(const oracle_problem *P, work_t *w) {
  const size_t KNp = (size_t)P->K * P->Np, KNF = (size_t)P->K * NFACE * P->Nfp, KF = (size_t)P->K * NFACE;
  memset(w, 0, sizeof(*w));
  w->maxv = malloc(KF * sizeof(double));
  w->maxv_fs = malloc(KF * sizeof(double));
  w->tau = malloc(KF * sizeof(double));
  w->ub = malloc(P->nvar * KNF * sizeof(double));
  w->fLM = malloc(P->nvar * KNF * sizeof(double));
  w->fR = malloc(P->nvar * KNp * sizeof(double));
  w->fS = malloc(P->nvar * KNp * sizeof(double));
  w->fT = malloc(P->nvar * KNp * sizeof(double));
  return (w->maxv && w->maxv_fs && w->tau && w->ub && w->fLM && w->fR && w->fS && w->fT) ? 0 : -1;
}

static void work_free(work_t *w) {
  free(w->maxv);
  free(w->maxv_fs);
  free(w->tau);
  free(w->ub);
  free(w->fLM);
  free(w->fR);
  free(w->fS);
  free(w->fT);
}

static void rhs_core(const oracle_problem *P, work_t *w, const double *u, double *rhs) {
  if (P->flux == 1) {
    calc_flux_rusanov(P, u, w->fLM);
  } else {
    update_ub(P, u, w->ub);
    calc_fluxLM(P, w->tau, w->ub, u, w->fLM);
  }
  calc_fluxes(P, u, w->fR, w->fS, w->fT);
  update_rhs(P, w->fR, w->fS, w->fT, w->fLM, rhs);
}

int oracle_rhs(const oracle_problem *P, const double *u, double *rhs, double *maxv, double *tau) {
  work_t w;
  if (work_alloc(P, &w)) {
    work_free(&w);
    return -1;
  }
  calc_maxv(P, u, w.maxv, w.maxv_fs);
  update_tau(P, w.maxv, w.tau);
  rhs_core(P, &w, u, rhs);
  const size_t KF = (size_t)P->K * NFACE;
  if (maxv) memcpy(maxv, w.maxv, KF * sizeof(double));
  if (tau) memcpy(tau, w.tau, KF * sizeof(double));
  work_free(&w);
  return 0;
}

int oracle_rhs_tau(const oracle_problem *P, const double *u, const double *tau, double *rhs) {
  work_t w;
  if (work_alloc(P, &w)) {
    work_free(&w);
    return -1;
  }
  memcpy(w.tau, tau, (size_t)P->K * NFACE * sizeof(double));
  rhs_core(P, &w, u, rhs);
  work_free(&w);
  return 0;
}

int oracle_run(const oracle_problem *P, double *u, double *k, int nsteps, double *t, int *tstep, double *dt_hist) {
  work_t w;
  const size_t n = (size_t)P->nvar * P->K * P->Np;
  double *utmp = malloc(n * sizeof(double));
  if (work_alloc(P, &w) || !utmp) {
    work_free(&w);
    free(utmp);
    return -1;
  }
  for (int st = 0; st < nsteps; st++) {
    if (!(*t < P->tmax)) {
      if (dt_hist) dt_hist[st] = 0.0;
      continue;
    }
    calc_maxv(P, u, w.maxv, w.maxv_fs);
    double dt = calc_dt(P, w.maxv_fs);
    if (*t + dt > P->tmax) dt = P->tmax - *t;
    update_tau(P, w.maxv, w.tau);
    for (int s = 0; s < 5; s++) {
      rhs_core(P, &w, u, utmp);
      const double a = RKA[s], b = RKB[s];
#pragma omp parallel for
      for (long i = 0; i < (long)n; i++) {
        k[i] = a * k[i] + dt * utmp[i];
        u[i] += b * k[i];
      }
    }
    if (dt_hist) dt_hist[st] = dt;
    *t += dt;
    *tstep += 1;
  }
  work_free(&w);
  free(utmp);
  return 0;
}

/* ------------------------------------------------------------------------
 * Tangent-linear RHS with tau frozen at u: drhs = dR/du(u) du.
 * J_n(U) dU of the Euler flux n.F (derivative of the flux formulas above);
 * pinned by central differences of oracle_rhs in tests/test_oracle.py.
 * ---------------------------------------------------------------------- */
static void euler_jv(const double U[5], double nx, double ny, double nz, const double d[5], double G, double out[5]) {
  double inv = 1.0 / U[0];
  double v[3] = {U[1] * inv, U[2] * inv, U[3] * inv}, n[3] = {nx, ny, nz};
  double q2 = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
  double p = (G - 1) * (U[4] - 0.5 * U[0] * q2);
  double H = (U[4] + p) * inv;
  double vn = n[0] * v[0] + n[1] * v[1] + n[2] * v[2];
  double gm = G - 1;
  double ndm = n[0] * d[1] + n[1] * d[2] + n[2] * d[3];
  double vdm = v[0] * d[1] + v[1] * d[2] + v[2] * d[3];
  out[0] = ndm;
  for (int c = 0; c < 3; c++)
    out[1 + c] = (-v[c] * vn + n[c] * gm * 0.5 * q2) * d[0] + v[c] * ndm + vn * d[1 + c] - n[c] * gm * vdm +
                 n[c] * gm * d[4];
  out[4] = vn * (gm * 0.5 * q2 - H) * d[0] + H * ndm - gm * vn * vdm + G * vn * d[4];
}

int oracle_rhs_tl(const oracle_problem *P, const double *u, const double *du, const double *tau, double *drhs) {
  const int K = P->K, Np = P->Np, Nfp = P->Nfp, NFF = NFACE * Nfp, nv = P->nvar;
  const size_t KNp = (size_t)K * Np, KNF = (size_t)K * NFF;
  if (P->flux != 0) return -2; /* the tangent-linear is of the LM flux only */
  work_t w;
  if (work_alloc(P, &w)) {
    work_free(&w);
    return -1;
  }
  if (tau) {
    memcpy(w.tau, tau, (size_t)K * NFACE * sizeof(double));
  } else {
    calc_maxv(P, u, w.maxv, w.maxv_fs);
    update_tau(P, w.maxv, w.tau);
  }
  /* dB volume -> w.fR/fS/fT, dB surface -> w.fLM */
#pragma omp parallel for
  for (int ei = 0; ei < K; ei++) {
    double m[9] = {P->rx[ei], P->ry[ei], P->rz[ei], P->sx[ei], P->sy[ei], P->sz[ei], P->tx[ei], P->ty[ei], P->tz[ei]};
    for (int j = 0; j < Np; j++) {
      size_t idx = (size_t)ei * Np + j;
      double JF[3][5];
      if (nv == 5) {
        double U[5], D[5];
        for (int v = 0; v < 5; v++) {
          U[v] = u[v * KNp + idx];
          D[v] = du[v * KNp + idx];
        }
        for (int x = 0; x < 3; x++) euler_jv(U, x == 0, x == 1, x == 2, D, P->gamma, JF[x]);
      } else {
        JF[0][0] = P->vx[idx] * du[idx];
        JF[1][0] = P->vy[idx] * du[idx];
        JF[2][0] = P->vz[idx] * du[idx];
      }
      for (int v = 0; v < nv; v++) {
        w.fR[v * KNp + idx] = m[0] * JF[0][v] + m[1] * JF[1][v] + m[2] * JF[2][v];
        w.fS[v * KNp + idx] = m[3] * JF[0][v] + m[4] * JF[1][v] + m[5] * JF[2][v];
        w.fT[v * KNp + idx] = m[6] * JF[0][v] + m[7] * JF[1][v] + m[8] * JF[2][v];
      }
    }
  }
#pragma omp parallel for
  for (int fidx = 0; fidx < K * NFACE; fidx++) {
    int ei = fidx / NFACE, fi = fidx % NFACE;
    double nx = P->nx[fidx], ny = P->ny[fidx], nz = P->nz[fidx], fs = P->Fscale[fidx], tau = w.tau[fidx];
    for (int j = 0; j < Nfp; j++) {
      size_t idx = (size_t)fidx * Nfp + j;
      size_t iM = (size_t)ei * Np + P->Fmask[fi * Nfp + j], iP = (size_t)P->vmapP[idx];
      /* a Dirichlet face's ub is data: no derivative */
      const int dir = P->gD && P->EtoB && P->EtoB[fidx] == 1;
      if (nv == 5) {
        double ub[5], dub[5], out[5];
        for (int v = 0; v < 5; v++) {
          ub[v] = dir ? P->gD[v * KNF + idx] : 0.5 * (u[v * KNp + iM] + u[v * KNp + iP]);
          dub[v] = dir ? 0.0 : 0.5 * (du[v * KNp + iM] + du[v * KNp + iP]);
        }
        euler_jv(ub, nx, ny, nz, dub, P->gamma, out);
        for (int v = 0; v < 5; v++) w.fLM[v * KNF + idx] = (out[v] - tau * (dub[v] - du[v * KNp + iM])) * fs;
      } else {
        double dub = dir ? 0.0 : 0.5 * (du[iM] + du[iP]);
        double vn = nx * P->vx[iM] + ny * P->vy[iM] + nz * P->vz[iM];
        w.fLM[idx] = (vn * dub - tau * (dub - du[iM])) * fs;
      }
    }
  }
  update_rhs(P, w.fR, w.fS, w.fT, w.fLM, drhs);
  work_free(&w);
  return 0;
}
