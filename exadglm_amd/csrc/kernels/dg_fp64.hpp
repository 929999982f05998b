// FP64 DGLM kernels: pointwise Euler fluxes and the fused per-stage
// RHS + LSRK kernel (volume flux, LM surface flux, stacked contraction with
// [WDr|WDs|WDt|-LIFT], LSRK register update) -- one launch per stage.
//
// Math follows the reference CPU path:
//   calc_maxv   dim3/src/cpp/euler_dglm.cpp:6-50,   advec_dglm.cpp:6-35
//               (its kernel, wavespeed_lds_kernel, is in dg_ws.hpp)
//   calc_dt     euler_dglm.cpp:53-69 (its kernel is in dg_step.hpp)
//   update_tau  euler_dglm.cpp:72-84
//   update_ub + calc_fluxLM  euler_dglm.cpp:87-183, advec_dglm.cpp:69-133
//   calc_fluxes euler_dglm.cpp:186-251, advec_dglm.cpp:136-169
//   update_rhs  euler_dglm.cpp:254-302
//   lserk_stage euler_time.cpp:15-38
#pragma once
#include "dg_common.hpp"

namespace dglm {

enum { PHYS_ADV = 0, PHYS_EULER = 1 };

template <int PHYS>
struct NVar {
  static constexpr int value = PHYS == PHYS_EULER ? 5 : 1;
};

// --------------------------------------------------------------------------
// Pointwise physics
// --------------------------------------------------------------------------

// Euler fluxes F (x), G (y), H (z) of the five conserved variables.
__device__ inline void euler_flux(const double u[5], double g, double F[5], double G[5], double H[5]) {
  const double inv = 1.0 / u[0];
  const double m2 = u[1] * u[1] + u[2] * u[2] + u[3] * u[3];
  const double p = (g - 1.0) * (u[4] - 0.5 * m2 * inv);
  const double vx = u[1] * inv, vy = u[2] * inv, vz = u[3] * inv;
  const double ep = u[4] + p;
  F[0] = u[1];
  F[1] = u[1] * vx + p;
  F[2] = u[2] * vx;
  F[3] = u[3] * vx;
  F[4] = ep * vx;
  G[0] = u[2];
  G[1] = u[1] * vy;
  G[2] = u[2] * vy + p;
  G[3] = u[3] * vy;
  G[4] = ep * vy;
  H[0] = u[3];
  H[1] = u[1] * vz;
  H[2] = u[2] * vz;
  H[3] = u[3] * vz + p;
  H[4] = ep * vz;
}

// Textbook Rusanov (local Lax-Friedrichs) normal flux of variable v:
//   F*.n = (n.F(u-) + n.F(u+))/2 + lambda/2 (u- - u+),
//   lambda = max(|v-.n| + c-, |v+.n| + c+)   (a flux option beside the LM flux).
__device__ inline double euler_normal_flux(int v, const double U[5], const double *n, double gamma, double &lam) {
  const double inv = 1.0 / U[0];
  const double p = (gamma - 1.0) * (U[4] - 0.5 * (U[1] * U[1] + U[2] * U[2] + U[3] * U[3]) * inv);
  const double vn = (n[0] * U[1] + n[1] * U[2] + n[2] * U[3]) * inv;
  lam = fabs(vn) + sqrt(fabs(gamma * p * inv));
  if (v == 0) return vn * U[0];
  if (v == 4) return (U[4] + p) * vn;
  return U[v] * vn + n[v - 1] * p;
}

__device__ inline double euler_rusanov(int v, const double uM[5], const double uP[5], const double *n, double gamma) {
  double lm, lp;
  const double fm = euler_normal_flux(v, uM, n, gamma, lm);
  const double fp = euler_normal_flux(v, uP, n, gamma, lp);
  return 0.5 * (fm + fp) + 0.5 * fmax(lm, lp) * (uM[v] - uP[v]);
}

// All five components of euler_rusanov with the primitives of each trace
// computed once (same arithmetic, so the same values).
__device__ inline void euler_normal_flux5(const double U[5], const double *n, double gamma, double F[5], double &lam) {
  const double inv = 1.0 / U[0];
  const double p = (gamma - 1.0) * (U[4] - 0.5 * (U[1] * U[1] + U[2] * U[2] + U[3] * U[3]) * inv);
  const double vn = (n[0] * U[1] + n[1] * U[2] + n[2] * U[3]) * inv;
  lam = fabs(vn) + sqrt(fabs(gamma * p * inv));
  F[0] = vn * U[0];
  F[1] = U[1] * vn + n[0] * p;
  F[2] = U[2] * vn + n[1] * p;
  F[3] = U[3] * vn + n[2] * p;
  F[4] = (U[4] + p) * vn;
}

__device__ inline void euler_rusanov5(const double uM[5], const double uP[5], const double *n, double gamma,
                                      double fl[5]) {
  double fm[5], fp[5], lm, lp;
  euler_normal_flux5(uM, n, gamma, fm, lm);
  euler_normal_flux5(uP, n, gamma, fp, lp);
  const double lam = fmax(lm, lp);
#pragma unroll
  for (int v = 0; v < 5; v++) fl[v] = 0.5 * (fm[v] + fp[v]) + 0.5 * lam * (uM[v] - uP[v]);
}

// --------------------------------------------------------------------------
// Fused stage kernel, straightforward LDS-staged form.
//   MODE 0: k = a k + dt rhs; u_out = u_in + b k
//   MODE 1: rhs only
// --------------------------------------------------------------------------
struct StageArgs {
  int K;
  const double *u_in;
  double *u_out;
  double *k;
  double *rhs;
  const double *vel;
  const double *geo;
  const int *conn;
  const int *fmx;
  const double *op;   // [KD][Np]: WDr^T, WDs^T, WDt^T, -LIFT^T
  const double *maxv;
  const DevScalars *scal;
  double rka, rkb;
  double gamma, tau_scale;
  int flux;      // 0: LM (reference), 1: Rusanov (upwind for advection)
  int bf_terms;  // split-BF16 contraction: 3 (two pieces per operand) or 6 (three pieces)
  // Dirichlet faces (EtoB == 1): dmap[e*4+f] = slot of the face (-1: not
  // Dirichlet), gbnd[v][slot][j] = prescribed state at its face node j (Fmask
  // order); nd slots.  dmap null: the mesh has none.
  const int *dmap;
  const double *gbnd;
  int nd;
};

// Prescribed boundary state of variable v at face node j of Dirichlet slot ds.
// The LM flux takes it as the face state ub (the reference leaves the branch
// as "ub = function of x,y,z,t", euler_dglm.cpp:136-141); the Rusanov option
// takes it as the exterior trace u+.
__device__ inline double dirichlet_value(const StageArgs &a, int v, int ds, int j, int Nfp) {
  return a.gbnd[((size_t)v * a.nd + ds) * Nfp + j];
}

template <int N, int PHYS>
constexpr int eb_simple() {
  using O = Ord<N>;
  constexpr int NV = NVar<PHYS>::value;
  constexpr int bytes =
      8 * (NV * (O::Np + O::KD) + (PHYS == PHYS_ADV ? 3 * O::Np : 0) + GEO_STRIDE + 4) + 16;
  constexpr int eb = 56000 / bytes;
  return eb < 1 ? 1 : (eb > 64 ? 64 : eb);
}

template <int N, int PHYS, int MODE>
__global__ __launch_bounds__(256) void stage_simple_kernel(StageArgs a) {
  using O = Ord<N>;
  constexpr int Np = O::Np, Nfp = O::Nfp, NFF = O::NFF, KD = O::KD;
  constexpr int NV = NVar<PHYS>::value;
  constexpr int EB = eb_simple<N, PHYS>();
  constexpr int NVEL = PHYS == PHYS_ADV ? 3 : 1;
  __shared__ double s_u[NV][EB * Np];
  __shared__ double s_f[NV][EB * KD];
  __shared__ double s_vel[NVEL][PHYS == PHYS_ADV ? EB * Np : 1];
  __shared__ double s_geo[EB * GEO_STRIDE];
  __shared__ double s_tau[EB * 4];
  __shared__ int s_conn[EB * 4];
  __shared__ int s_dm[EB * 4];
  __shared__ int s_fmx[16 * Nfp];

  const int K = a.K;
  const size_t KNp = (size_t)K * Np;
  const int e0 = blockIdx.x * EB;
  const int ne = min(EB, K - e0);
  const int tid = threadIdx.x;

  if (MODE == 0 && !a.scal->active) {
    // A step past tmax: keep the state (the caller's buffer swap still happens).
    for (int v = 0; v < NV; v++)
      for (int i = tid; i < ne * Np; i += 256) a.u_out[v * KNp + (size_t)e0 * Np + i] = a.u_in[v * KNp + (size_t)e0 * Np + i];
    return;
  }

  for (int i = tid; i < 16 * Nfp; i += 256) s_fmx[i] = a.fmx[i];
  for (int i = tid; i < ne * GEO_STRIDE; i += 256) s_geo[i] = a.geo[(size_t)e0 * GEO_STRIDE + i];
  for (int i = tid; i < ne * 4; i += 256) {
    const int c = a.conn[(size_t)e0 * 4 + i];
    s_conn[i] = c;
    const int nf = conn_nbr(c) * 4 + conn_face(c);
    s_tau[i] = 0.25 * (a.maxv[(size_t)e0 * 4 + i] + a.maxv[nf]) * a.tau_scale;
    s_dm[i] = a.dmap ? a.dmap[(size_t)e0 * 4 + i] : -1;
  }
  for (int v = 0; v < NV; v++)
    for (int i = tid; i < ne * Np; i += 256) s_u[v][i] = a.u_in[v * KNp + (size_t)e0 * Np + i];
  if (PHYS == PHYS_ADV)
    for (int d = 0; d < 3; d++)
      for (int i = tid; i < ne * Np; i += 256) s_vel[d][i] = a.vel[d * KNp + (size_t)e0 * Np + i];
  __syncthreads();

  // Volume fluxes: fR = rx F + ry G + rz H, fS, fT likewise.
  for (int idx = tid; idx < ne * Np; idx += 256) {
    const int e = idx / Np, i = idx - e * Np;
    const double *g = &s_geo[e * GEO_STRIDE];
    if (PHYS == PHYS_EULER) {
      double u[5], F[5], G[5], H[5];
      for (int v = 0; v < 5; v++) u[v] = s_u[v][idx];
      euler_flux(u, a.gamma, F, G, H);
      for (int v = 0; v < 5; v++) {
        double *fo = &s_f[v][e * KD];
        fo[i] = g[0] * F[v] + g[1] * G[v] + g[2] * H[v];
        fo[Np + i] = g[3] * F[v] + g[4] * G[v] + g[5] * H[v];
        fo[2 * Np + i] = g[6] * F[v] + g[7] * G[v] + g[8] * H[v];
      }
    } else {
      const double uu = s_u[0][idx];
      const double F = s_vel[0][idx] * uu, G = s_vel[1][idx] * uu, H = s_vel[2][idx] * uu;
      double *fo = &s_f[0][e * KD];
      fo[i] = g[0] * F + g[1] * G + g[2] * H;
      fo[Np + i] = g[3] * F + g[4] * G + g[5] * H;
      fo[2 * Np + i] = g[6] * F + g[7] * G + g[8] * H;
    }
  }

  // LM surface flux (n.F(ub) - tau (ub - u-)) Fscale with ub = (u- + u+)/2,
  // or ub = the prescribed state on Dirichlet faces (Rusanov: u+ = it).
  for (int idx = tid; idx < ne * NFF; idx += 256) {
    const int e = idx / NFF, m = idx - e * NFF;
    const int f = m / Nfp, j = m - f * Nfp;
    const int c = s_conn[e * 4 + f];
    const size_t pidx = (size_t)conn_nbr(c) * Np + s_fmx[(conn_face(c) * 4 + conn_code(c)) * Nfp + j];
    const int mloc = e * Np + s_fmx[(f * 4 + 3) * Nfp + j];
    const double *g = &s_geo[e * GEO_STRIDE + 9 + f * 4];
    const double nx = g[0], ny = g[1], nz = g[2], fs = g[3];
    const double tau = s_tau[e * 4 + f];
    const int ds = s_dm[e * 4 + f];
    if (PHYS == PHYS_EULER) {
      double uM[5], uP[5], ub[5], F[5], G[5], H[5];
      for (int v = 0; v < 5; v++) {
        uM[v] = s_u[v][mloc];
        uP[v] = ds >= 0 ? dirichlet_value(a, v, ds, j, Nfp) : a.u_in[v * KNp + pidx];
        ub[v] = ds >= 0 ? uP[v] : 0.5 * (uM[v] + uP[v]);
      }
      if (a.flux == 1) {
        for (int v = 0; v < 5; v++) s_f[v][e * KD + 3 * Np + m] = euler_rusanov(v, uM, uP, g, a.gamma) * fs;
      } else {
        euler_flux(ub, a.gamma, F, G, H);
        for (int v = 0; v < 5; v++)
          s_f[v][e * KD + 3 * Np + m] = ((nx * F[v] + ny * G[v] + nz * H[v]) - tau * (ub[v] - uM[v])) * fs;
      }
    } else {
      const double uM = s_u[0][mloc];
      const double uP = ds >= 0 ? dirichlet_value(a, 0, ds, j, Nfp) : a.u_in[pidx];
      const double ub = ds >= 0 && a.flux != 1 ? uP : 0.5 * (uM + uP);
      const double vn = nx * s_vel[0][mloc] + ny * s_vel[1][mloc] + nz * s_vel[2][mloc];
      if (a.flux == 1) {
        s_f[0][e * KD + 3 * Np + m] = (vn * ub + 0.5 * fabs(vn) * (uM - uP)) * fs;
      } else {
        const double nf = nx * (s_vel[0][mloc] * ub) + ny * (s_vel[1][mloc] * ub) + nz * (s_vel[2][mloc] * ub);
        s_f[0][e * KD + 3 * Np + m] = (nf - tau * (ub - uM)) * fs;
      }
    }
  }
  __syncthreads();

  // Contraction rhs_i = sum_k op[k][i] f[k], then the LSRK register update.
  const double dt = MODE == 0 ? a.scal->dt : 0.0;
  for (int o = tid; o < NV * ne * Np; o += 256) {
    const int v = o / (ne * Np), r = o - v * (ne * Np);
    const int e = r / Np, i = r - e * Np;
    const double *F = &s_f[v][e * KD];
    double acc = 0.0;
#pragma unroll 5
    for (int kk = 0; kk < KD; kk++) acc += a.op[kk * Np + i] * F[kk];
    const size_t gi = v * KNp + (size_t)e0 * Np + r;
    if (MODE == 1) {
      a.rhs[gi] = acc;
    } else {
      const double kn = a.rka * a.k[gi] + dt * acc;
      a.k[gi] = kn;
      a.u_out[gi] = s_u[v][r] + a.rkb * kn;
    }
  }
}

}  // namespace dglm
