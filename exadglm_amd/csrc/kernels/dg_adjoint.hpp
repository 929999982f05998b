// Discrete adjoint of the DGLM RHS and of one LSRK4(5) stage (FP64).
//
// Forward stage (see dg_fp64.hpp):  r = R(u; tau) = A B(u; tau),
//   A = [WDr | WDs | WDt | -LIFT],  B = stacked volume fluxes and LM surface
//   fluxes;  k' = a k + dt r;  u' = u + b k'.
// Reverse stage, given the adjoints (ubar, kbar) of (u', k'):
//   kb = kbar + b ubar;  rbar = dt kb;  kbar <- a kb;
//   ubar <- ubar + J_B(u)^T (A^T rbar),   taubar_f += dB/dtau_f . (A^T rbar)
// J_B^T has three parts:
//   volume   B[d*Np+j] = sum_x metric[d][x] F^x(u_j): pointwise transposed
//            Jacobians of F, G, H at node j;
//   own face fLM(m) = (n.F(ub) - tau (ub - u-)) Fscale, ub = (u- + u+)/2:
//            d/du- = Fscale (J_n(ub)/2 + tau/2), added at the face node's
//            volume node;
//   partner  the same flux seen from the neighbour element depends on this
//            element's node through its u+:  d/du+ = Fscale_nbr (J_nbr(ub)/2
//            - tau/2).  The neighbour's pass A forms it beside its own-side
//            term (the same ub, normal, h and tau) and stores it at its face
//            slot; each element PULLS these terms through the partner slot
//            (exact by the involution of the face connectivity) -- no
//            atomics, so the adjoint is bitwise deterministic.
// Pass A computes A^T rbar, the volume and own-face terms and tau-bar, and
// stores the partner-side face terms; pass B adds them.
#pragma once
#include "dg_fp64.hpp"

namespace dglm {

// w += J_n(U)^T h, J_n = d(n.F)/dU of the Euler flux at U (5 conserved vars).
__device__ inline void euler_jt(const double U[5], double nx, double ny, double nz, const double h[5], double g,
                                double w[5]) {
  const double inv = 1.0 / U[0];
  const double v[3] = {U[1] * inv, U[2] * inv, U[3] * inv};
  const double n[3] = {nx, ny, nz};
  const double q2 = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
  const double p = (g - 1.0) * (U[4] - 0.5 * U[0] * q2);
  const double H = (U[4] + p) * inv;
  const double vn = nx * v[0] + ny * v[1] + nz * v[2];
  const double gm = g - 1.0;
  // rho column
  double w0 = h[4] * vn * (gm * 0.5 * q2 - H);
  double hn = 0.0, hv = 0.0;  // sum_c h_c n_c, sum_c h_c v_c
  for (int c = 0; c < 3; c++) {
    w0 += h[1 + c] * (-v[c] * vn + n[c] * gm * 0.5 * q2);
    hn += h[1 + c] * n[c];
    hv += h[1 + c] * v[c];
  }
  w[0] += w0;
  // momentum columns j: h0 n_j + sum_c h_c (v_c n_j + d_cj vn - n_c gm v_j) + h4 (H n_j - gm v_j vn)
  for (int j = 0; j < 3; j++)
    w[1 + j] += h[0] * n[j] + hv * n[j] + h[1 + j] * vn - hn * gm * v[j] + h[4] * (H * n[j] - gm * v[j] * vn);
  // energy column
  w[4] += hn * gm + h[4] * g * vn;
}

struct AdjArgs {
  int K;
  int mode;            // 0: reverse LSRK stage; 1: plain RHS adjoint (rbar given, ubar := J^T rbar)
  const double *u;     // forward stage state u_s
  double *ubar;        // [NV][K*Np]  in/out
  double *kbar;        // [NV][K*Np]  in/out (mode 0)
  const double *rbar;  // [NV][K*Np]  (mode 1)
  double *gsur;        // [K][NFF][NV] partner-side face terms (pass A out, pass B in)
  // partner-side face terms of the previous pass A, not yet added to ubar by
  // a pass B: the MFMA pass A adds them to ubar itself (null: none pending)
  const double *gsur_in;
  double *taubar;      // [K*4] accumulated tau adjoint per face side (may be null)
  const double *vel;
  const double *geo;
  const int *conn;
  const int *fmx;
  const int *node_slots;  // [Np][3] face-node slots m = f*Nfp + j with Fmask[f][j] == node (-1 pad)
  const int *fpos;        // [4][Np] position of a volume node in Fmask[f] (-1 if not on face f)
  const double *op;       // [KD][Np]
  const double *maxv;     // wave speed per face at the step start (tau)
  const double *dtp;      // device dt of the step (mode 0)
  // dt sensitivity (exact adjoint): r_s is recovered from the stage states,
  // k_{s+1} = (u_{s+1} - u_s)/b_s, r_s = (k_{s+1} - a_s k_s)/dt, and
  // dtpart[block] += sum_nodes r_s . kb
  const double *u_next, *u_prev;
  double *dtpart;
  double a, b, a_prev, b_prev;
  double gamma, tau_scale;
};

template <int N, int PHYS>
constexpr int eb_adjoint() {
  using O = Ord<N>;
  constexpr int NV = NVar<PHYS>::value;
  constexpr int bytes = 8 * (NV * (3 * O::Np + O::KD) + (PHYS == PHYS_ADV ? 3 * O::Np : 0) + GEO_STRIDE + 8) + 16;
  constexpr int eb = 56000 / bytes;
  return eb < 1 ? 1 : (eb > 64 ? 64 : eb);
}

// Pass A.
template <int N, int PHYS>
__global__ __launch_bounds__(256) void adjoint_pass_a(AdjArgs a) {
  using O = Ord<N>;
  constexpr int Np = O::Np, Nfp = O::Nfp, NFF = O::NFF, KD = O::KD;
  constexpr int NV = NVar<PHYS>::value;
  constexpr int EB = eb_adjoint<N, PHYS>();
  constexpr int NVEL = PHYS == PHYS_ADV ? 3 : 1;
  __shared__ double s_u[NV][EB * Np];
  __shared__ double s_r[NV][EB * Np];   // rbar
  __shared__ double s_w[NV][EB * Np];   // ubar accumulator
  __shared__ double s_g[NV][EB * KD];   // A^T rbar
  __shared__ double s_vel[NVEL][PHYS == PHYS_ADV ? EB * Np : 1];
  __shared__ double s_geo[EB * GEO_STRIDE];
  __shared__ double s_tau[EB * 4];
  __shared__ int s_conn[EB * 4];
  __shared__ int s_fmx[16 * Nfp];
  __shared__ int s_slots[Np * 3];

  const int K = a.K;
  const size_t KNp = (size_t)K * Np;
  const int e0 = blockIdx.x * EB;
  const int ne = min(EB, K - e0);
  const int tid = threadIdx.x;

  for (int i = tid; i < 16 * Nfp; i += 256) s_fmx[i] = a.fmx[i];
  for (int i = tid; i < Np * 3; i += 256) s_slots[i] = a.node_slots[i];
  for (int i = tid; i < ne * GEO_STRIDE; i += 256) s_geo[i] = a.geo[(size_t)e0 * GEO_STRIDE + i];
  for (int i = tid; i < ne * 4; i += 256) {
    const int c = a.conn[(size_t)e0 * 4 + i];
    s_conn[i] = c;
    s_tau[i] = 0.25 * (a.maxv[(size_t)e0 * 4 + i] + a.maxv[conn_nbr(c) * 4 + conn_face(c)]) * a.tau_scale;
  }
  const double dt = a.mode == 0 ? *a.dtp : 0.0;
  double dtsum = 0.0;
  for (int v = 0; v < NV; v++)
    for (int i = tid; i < ne * Np; i += 256) {
      const size_t gi = v * KNp + (size_t)e0 * Np + i;
      const double uu = a.u[gi];
      s_u[v][i] = uu;
      if (a.mode == 0) {
        const double ub = a.ubar[gi];
        const double kb = a.kbar[gi] + a.b * ub;
        a.kbar[gi] = a.a * kb;
        s_r[v][i] = dt * kb;
        s_w[v][i] = ub;
        if (a.dtpart && dt != 0.0) {
          const double knext = (a.u_next[gi] - uu) / a.b;
          const double kcur = a.u_prev ? (uu - a.u_prev[gi]) / a.b_prev : 0.0;
          dtsum += (knext - a.a * kcur) / dt * kb;
        }
      } else {
        s_r[v][i] = a.rbar[gi];
        s_w[v][i] = 0.0;
      }
    }
  if (a.mode == 0 && a.dtpart) {
    // deterministic block sum into this block's slot
    __shared__ double red[256];
    red[tid] = dtsum;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
      if (tid < w) red[tid] += red[tid + w];
      __syncthreads();
    }
    if (tid == 0) a.dtpart[blockIdx.x] += red[0];
  }
  if (PHYS == PHYS_ADV)
    for (int d = 0; d < 3; d++)
      for (int i = tid; i < ne * Np; i += 256) s_vel[d][i] = a.vel[d * KNp + (size_t)e0 * Np + i];
  __syncthreads();

  // g = A^T rbar : g[k] = sum_i op[k][i] rbar[i]
  for (int idx = tid; idx < ne * KD; idx += 256) {
    const int e = idx / KD, k = idx - e * KD;
    double acc[NV];
    for (int v = 0; v < NV; v++) acc[v] = 0.0;
    const double *opk = a.op + (size_t)k * Np;
    for (int i = 0; i < Np; i++) {
      const double o = opk[i];
      for (int v = 0; v < NV; v++) acc[v] += o * s_r[v][e * Np + i];
    }
    for (int v = 0; v < NV; v++) s_g[v][idx] = acc[v];
  }
  __syncthreads();

  // Volume part: ubar(j) += sum_x J_x(u_j)^T h^x,  h^x_v = sum_d metric[d][x] g[d*Np+j][v].
  for (int idx = tid; idx < ne * Np; idx += 256) {
    const int e = idx / Np, j = idx - e * Np;
    const double *m = &s_geo[e * GEO_STRIDE];
    if (PHYS == PHYS_EULER) {
      double U[5], w[5] = {0, 0, 0, 0, 0};
      for (int v = 0; v < 5; v++) U[v] = s_u[v][idx];
      for (int x = 0; x < 3; x++) {
        double h[5];
        for (int v = 0; v < 5; v++) {
          const double *gv = &s_g[v][e * KD];
          h[v] = m[0 + x] * gv[j] + m[3 + x] * gv[Np + j] + m[6 + x] * gv[2 * Np + j];
        }
        euler_jt(U, x == 0 ? 1.0 : 0.0, x == 1 ? 1.0 : 0.0, x == 2 ? 1.0 : 0.0, h, a.gamma, w);
      }
      for (int v = 0; v < 5; v++) s_w[v][idx] += w[v];
    } else {
      const double *gv = &s_g[0][e * KD];
      double w = 0.0;
      for (int x = 0; x < 3; x++)
        w += s_vel[x][idx] * (m[0 + x] * gv[j] + m[3 + x] * gv[Np + j] + m[6 + x] * gv[2 * Np + j]);
      s_w[0][idx] += w;
    }
  }
  __syncthreads();

  // Own-face part, gathered per volume node over its (at most 3) face slots.
  for (int idx = tid; idx < ne * Np; idx += 256) {
    const int e = idx / Np, i = idx - e * Np;
    for (int q = 0; q < 3; q++) {
      const int mslot = s_slots[i * 3 + q];
      if (mslot < 0) break;
      const int f = mslot / Nfp, j = mslot - f * Nfp;
      const int c = s_conn[e * 4 + f];
      const size_t pidx = (size_t)conn_nbr(c) * Np + s_fmx[(conn_face(c) * 4 + conn_code(c)) * Nfp + j];
      const double *g = &s_geo[e * GEO_STRIDE + 9 + f * 4];
      const double tau = s_tau[e * 4 + f];
      if (PHYS == PHYS_EULER) {
        double ub[5], h[5], w[5] = {0, 0, 0, 0, 0};
        for (int v = 0; v < 5; v++) {
          ub[v] = 0.5 * (s_u[v][idx] + a.u[v * KNp + pidx]);
          h[v] = 0.5 * g[3] * s_g[v][e * KD + 3 * Np + mslot];
        }
        euler_jt(ub, g[0], g[1], g[2], h, a.gamma, w);
        // own side: (1/2 n.J + tau/2)^T Fscale g; the partner's node gets
        // (1/2 n.J - tau/2)^T Fscale g, the same terms, stored for pass B
        double *pt = &a.gsur[((size_t)(e0 + e) * NFF + mslot) * NV];
        for (int v = 0; v < 5; v++) {
          s_w[v][idx] += w[v] + tau * h[v];
          pt[v] = w[v] - tau * h[v];
        }
      } else {
        const double gg = s_g[0][e * KD + 3 * Np + mslot];
        const double vn = g[0] * s_vel[0][idx] + g[1] * s_vel[1][idx] + g[2] * s_vel[2][idx];
        s_w[0][idx] += g[3] * (0.5 * vn + 0.5 * tau) * gg;
        a.gsur[(size_t)(e0 + e) * NFF + mslot] = g[3] * (0.5 * vn - 0.5 * tau) * gg;
      }
    }
  }
  // tau adjoint per face side: sum_m sum_v g[m][v] * (-(ub_v - u-_v) Fscale)
  if (a.taubar)
    for (int idx = tid; idx < ne * 4; idx += 256) {
      const int e = idx >> 2, f = idx & 3;
      const int c = s_conn[idx];
      const double fs = s_geo[e * GEO_STRIDE + 9 + f * 4 + 3];
      double t = 0.0;
      for (int j = 0; j < Nfp; j++) {
        const int mloc = e * Np + s_fmx[(f * 4 + 3) * Nfp + j];
        const size_t pidx = (size_t)conn_nbr(c) * Np + s_fmx[(conn_face(c) * 4 + conn_code(c)) * Nfp + j];
        for (int v = 0; v < NV; v++) {
          const double um = s_u[v][mloc];
          const double ub = 0.5 * (um + a.u[v * KNp + pidx]);
          t += s_g[v][e * KD + 3 * Np + f * Nfp + j] * (-(ub - um) * fs);
        }
      }
      const size_t o = (size_t)e0 * 4 + idx;
      a.taubar[o] = (a.mode == 0 ? a.taubar[o] : 0.0) + t;
    }
  __syncthreads();

  for (int v = 0; v < NV; v++)
    for (int i = tid; i < ne * Np; i += 256) a.ubar[v * KNp + (size_t)e0 * Np + i] = s_w[v][i];
}

// Pass B: one thread per volume node; pull the partner-side terms of the
// (at most 3) faces the node lies on.  The partner element's pass A formed
// them (same ub, normal, h and tau as its own-side term) and stored them at
// its face-node slot, so only the connectivity, that slot and ubar are read.
template <int N, int PHYS>
__global__ __launch_bounds__(256) void adjoint_pass_b(AdjArgs a) {
  using O = Ord<N>;
  constexpr int Np = O::Np, Nfp = O::Nfp, NFF = O::NFF;
  constexpr int NV = NVar<PHYS>::value;
  __shared__ int s_slots[Np * 3], s_fmx[16 * Nfp], s_fpos[4 * Np];
  for (int i = threadIdx.x; i < Np * 3; i += 256) s_slots[i] = a.node_slots[i];
  for (int i = threadIdx.x; i < 16 * Nfp; i += 256) s_fmx[i] = a.fmx[i];
  for (int i = threadIdx.x; i < 4 * Np; i += 256) s_fpos[i] = a.fpos[i];
  __syncthreads();
  const int K = a.K;
  const size_t KNp = (size_t)K * Np;
  const int idx = blockIdx.x * 256 + threadIdx.x;
  if (idx >= K * Np) return;
  const int e = idx / Np, i = idx - e * Np;
  // all loads of the (at most 3) partner slots first, then the sums
  double t[3][NV];
#pragma unroll
  for (int q = 0; q < 3; q++) {
    const int ms = s_slots[i * 3 + q];
#pragma unroll
    for (int v = 0; v < NV; v++) t[q][v] = 0.0;
    if (ms >= 0) {
      const int f = ms / Nfp, j = ms - f * Nfp;
      const int c = a.conn[(size_t)e * 4 + f];
      const int nb = conn_nbr(c), nf = conn_face(c);
      const int pnode = s_fmx[(nf * 4 + conn_code(c)) * Nfp + j];  // neighbour's volume node
      const int jp = s_fpos[nf * Np + pnode];                       // its slot on face nf
      const double *gs = &a.gsur[((size_t)nb * NFF + nf * Nfp + jp) * NV];
#pragma unroll
      for (int v = 0; v < NV; v++) t[q][v] = gs[v];
    }
  }
  for (int v = 0; v < NV; v++) {
    double w = 0.0;
#pragma unroll
    for (int q = 0; q < 3; q++) w += t[q][v];
    a.ubar[v * KNp + idx] += w;
  }
}

// dtbar of the step = sum of the block partials + tbar (adjoint of the time
// after the step).  If the step was clipped to end at tmax, dt = tmax - t
// does not depend on the wave speed and its adjoint flows into t instead.
// One block; scal[0] = tbar (in/out), scal[1] = dtbar seen by lambda (out).
__global__ __launch_bounds__(256) void adjoint_dt_reduce(const double *__restrict__ dtpart, int n,
                                                         const int *__restrict__ clipped, int step, double *scal) {
  __shared__ double red[256];
  double s = 0.0;
  for (int i = threadIdx.x; i < n; i += 256) s += dtpart[i];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const double dtbar = red[0] + scal[0];
    if (clipped[step]) {
      scal[0] -= dtbar;
      scal[1] = 0.0;
    } else {
      scal[1] = dtbar;
    }
  }
}

// Wave-speed adjoint of one step (Euler): lambda_f enters tau of both sides
// of face f and, for the face attaining max(lambda*Fscale), dt.  The
// derivative of lambda = sqrt(max v^2) + sqrt(max |gamma p/rho|) follows the
// argmax nodes the forward picked (first strict maximum, as calc_maxv).
// A block takes LAM_ADJ_EB(Np) elements: |v|^2 and |gamma p/rho| of their
// nodes are formed with coalesced loads into LDS (the arithmetic of the
// forward wave-speed kernel, so the same argmax), then one thread per face side
// scans its face nodes and forms the ubar increments at the argmax nodes into
// LDS, and four threads per element (one or two variables each) add them in
// face order, so the accumulation order per node is fixed and the result
// deterministic.
__host__ __device__ constexpr int lam_adj_elems(int Np) { return Np <= 35 ? 64 : (Np <= 84 ? 32 : 16); }

template <int N>
__global__ __launch_bounds__(4 * lam_adj_elems(Ord<N>::Np)) void lambda_adjoint_kernel(
    int K, const double *__restrict__ u, const double *__restrict__ lam, const double *__restrict__ taubar,
    const int *__restrict__ conn, const int *__restrict__ fmx, const double *dt_seq, const int *fstar_seq, int step,
    const double *scal, double tau_scale, double gamma, double *ubar) {
  using O = Ord<N>;
  constexpr int Np = O::Np, Nfp = O::Nfp;
  constexpr int EB = lam_adj_elems(Np), NT = 4 * EB;
  __shared__ double s_v2[EB * Np], s_c2[EB * Np];
  __shared__ int s_fm[4 * Nfp];
  __shared__ int s_nv[NT], s_nc[NT];  // argmax volume nodes (-1: none)
  __shared__ double s_dv[4][NT];      // increments of rho, rho v at the |v|^2 argmax
  __shared__ double s_dc[5][NT];      // increments of all five at the c^2 argmax
  const size_t KNp = (size_t)K * Np;
  const int t = threadIdx.x;
  const int e0 = blockIdx.x * EB;
  const int ne = min(EB, K - e0);
  for (int i = t; i < 4 * Nfp; i += NT) s_fm[i] = fmx[((i / Nfp) * 4 + 3) * Nfp + i % Nfp];
  for (int i = t; i < ne * Np; i += NT) {
    const size_t idx = (size_t)e0 * Np + i;
    const double u1 = u[idx], u2 = u[KNp + idx], u3 = u[2 * KNp + idx], u4 = u[3 * KNp + idx],
                 u5 = u[4 * KNp + idx];
    const double vx = u2 / u1, vy = u3 / u1, vz = u4 / u1;
    const double v2 = vx * vx + vy * vy + vz * vz;
    const double pres = (gamma - 1.0) * (u5 - 0.5 * u1 * v2);
    s_v2[i] = v2;
    s_c2[i] = fabs(gamma * pres / u1);
  }
  __syncthreads();
  const int el = t >> 2, f = t & 3, e = e0 + el;
  int nv = -1, nc = -1;
  if (el < ne) {
    const double dtbar = scal[1], dt = dt_seq[step];
    const int o = e * 4 + f;
    const int c = conn[o];
    const int nbf = conn_nbr(c) * 4 + conn_face(c);
    double lb = 0.25 * tau_scale * (taubar[o] + taubar[nbf]);
    if (o == fstar_seq[step]) lb += -dtbar * dt / lam[o];
    if (lb != 0.0) {
      double mv2 = 0.0, mc2 = 0.0;
      int jv = -1, jc = -1;
      for (int j = 0; j < Nfp; j++) {
        const int q = el * Np + s_fm[f * Nfp + j];
        const double v2 = s_v2[q], c2 = s_c2[q];
        if (v2 > mv2) {
          mv2 = v2;
          jv = j;
        }
        if (c2 > mc2) {
          mc2 = c2;
          jc = j;
        }
      }
      if (jv >= 0) {
        nv = s_fm[f * Nfp + jv];
        const size_t idx = (size_t)e * Np + nv;
        const double r = u[idx], v[3] = {u[KNp + idx] / r, u[2 * KNp + idx] / r, u[3 * KNp + idx] / r};
        const double s = lb / (2.0 * sqrt(mv2));
        s_dv[0][t] = s * (-2.0 * mv2 / r);
        for (int q = 0; q < 3; q++) s_dv[1 + q][t] = s * (2.0 * v[q] / r);
      }
      if (jc >= 0) {
        nc = s_fm[f * Nfp + jc];
        const size_t idx = (size_t)e * Np + nc;
        const double r = u[idx], v[3] = {u[KNp + idx] / r, u[2 * KNp + idx] / r, u[3 * KNp + idx] / r};
        const double v2 = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
        const double gm = gamma - 1.0;
        const double p = gm * (u[4 * KNp + idx] - 0.5 * r * v2);
        const double sg = gamma * p / r >= 0.0 ? 1.0 : -1.0;
        const double s = sg * lb / (2.0 * sqrt(mc2));
        s_dc[0][t] = s * gamma * (gm * 0.5 * v2 / r - p / (r * r));
        for (int q = 0; q < 3; q++) s_dc[1 + q][t] = s * gamma * (-gm * v[q]) / r;
        s_dc[4][t] = s * gamma * gm / r;
      }
    }
  }
  s_nv[t] = nv;
  s_nc[t] = nc;
  __syncthreads();
  // thread (element el, k) adds variable k (and variable 4 for k = 0) over
  // the element's four faces in face order; distinct variables never alias
  if (el < ne) {
    const int k = f;
#pragma unroll
    for (int x = 0; x < 2; x++) {
      const int v = x == 0 ? k : 4;
      if (x == 1 && k != 0) break;
      for (int ff = 0; ff < 4; ff++) {
        const int q = 4 * el + ff;
        if (v < 4 && s_nv[q] >= 0) ubar[v * KNp + (size_t)e * Np + s_nv[q]] += s_dv[v][q];
        if (s_nc[q] >= 0) ubar[v * KNp + (size_t)e * Np + s_nc[q]] += s_dc[v][q];
      }
    }
  }
}

}  // namespace dglm
