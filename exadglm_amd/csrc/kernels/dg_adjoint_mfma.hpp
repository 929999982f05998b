// Pass A of the discrete Euler adjoint with the contraction g = A^T rbar on
// FP64 MFMA (gfx950).  Same mathematics as adjoint_pass_a (dg_adjoint.hpp);
// the difference is the layout of the work:
//
//   * a workgroup takes 16 elements (the MFMA columns) and has one wave per
//     16-row tile of A^T (KT = ceil(KD/16) waves, 11 at N=4); wave T keeps
//     the A^T fragments of its rows for the k-steps over the Np nodes in
//     registers and runs ceil(Np/4) v_mfma_f64_16x16x4f64 per variable;
//   * phase 0 stages u, forms rbar = dt (kbar + b ubar) (and updates kbar,
//     the dt partial sum) into an LDS tile [node][element];
//   * phase 1 is the contraction, g[v][k][e] into LDS (all five variables,
//     since the transposed flux Jacobians mix them);
//   * phase 2 is pointwise per volume node: the transposed volume Jacobians,
//     the node's own-face terms (exterior trace gathered from global memory)
//     and the write of ubar; then tau-bar per face side and the face part of
//     g for pass B.
#pragma once
#include "dg_adjoint.hpp"
#include "dg_mfma.hpp"

namespace dglm {

template <int N>
struct AdjShape {
  using O = Ord<N>;
  static constexpr int Np = O::Np, Nfp = O::Nfp, NFF = O::NFF, KD = O::KD;
  static constexpr int KT = (KD + 15) / 16;  // row tiles of A^T = waves
  static constexpr int KSA = (Np + 3) / 4;   // k-steps over the nodes
  static constexpr int NpA = 4 * KSA;        // padded node count of the rbar tile
  static constexpr int NT = 64 * KT;
  static constexpr int EB = 16;
};

template <int N>
constexpr size_t adj_mfma_lds_bytes() {
  using S = AdjShape<N>;
  return 8 * (5 * S::EB * S::Np + 5 * S::NpA * S::EB + 5 * S::KD * S::EB + S::EB * GEO_STRIDE + S::EB * 4 +
              S::NT / 64) +
         4 * (S::EB * 4 + 16 * S::Nfp + S::Np * 3);
}

template <int N>
constexpr bool adj_mfma_available() {
  using S = AdjShape<N>;
  return S::NT <= 1024 && adj_mfma_lds_bytes<N>() <= 160 * 1024;
}

// Host: afragT[(T*KSA + ks)*64 + l] = A^T[16T + l%16][4ks + l/16]
//                                   = A[4ks + l/16][16T + l%16] = opT[k*Np + i].
template <int N>
inline void adj_arrange(const double *opT, double *afragT) {
  using S = AdjShape<N>;
  for (int T = 0; T < S::KT; T++)
    for (int ks = 0; ks < S::KSA; ks++)
      for (int l = 0; l < 64; l++) {
        const int k = 16 * T + (l & 15), i = 4 * ks + (l >> 4);
        afragT[((size_t)T * S::KSA + ks) * 64 + l] = (k < S::KD && i < S::Np) ? opT[(size_t)k * S::Np + i] : 0.0;
      }
}

template <int N>
__global__ __launch_bounds__(AdjShape<N>::NT) void adjoint_pass_a_mfma(AdjArgs a, const double *__restrict__ afragT) {
  using S = AdjShape<N>;
  using P = PrecTraits<PREC_F64>;
  constexpr int Np = S::Np, Nfp = S::Nfp, NFF = S::NFF, KD = S::KD, KSA = S::KSA, NpA = S::NpA, NT = S::NT;
  constexpr int EB = S::EB;
  __shared__ double s_u[5][EB * Np];   // [element][node]
  __shared__ double s_r[5][NpA * EB];  // rbar, [node][element]
  __shared__ double s_g[5][KD * EB];   // A^T rbar, [row][element]
  __shared__ double s_geo[EB * GEO_STRIDE];
  __shared__ double s_tau[EB * 4];
  __shared__ int s_conn[EB * 4];
  __shared__ int s_fmx[16 * Nfp];
  __shared__ int s_slots[Np * 3];
  __shared__ double red[NT / 64];

  const int K = a.K;
  const size_t KNp = (size_t)K * Np;
  const int e0 = blockIdx.x * EB;
  const int ne = min(EB, K - e0);
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int w = __builtin_amdgcn_readfirstlane(tid >> 6);

  // ---- phase 0: staging, rbar, kbar update, dt partial ----
  for (int i = tid; i < 16 * Nfp; i += NT) s_fmx[i] = a.fmx[i];
  for (int i = tid; i < Np * 3; i += NT) s_slots[i] = a.node_slots[i];
  for (int i = tid; i < ne * GEO_STRIDE; i += NT) s_geo[i] = a.geo[(size_t)e0 * GEO_STRIDE + i];
  for (int i = tid; i < ne * 4; i += NT) {
    const int c = a.conn[(size_t)e0 * 4 + i];
    s_conn[i] = c;
    s_tau[i] = 0.25 * (a.maxv[(size_t)e0 * 4 + i] + a.maxv[conn_nbr(c) * 4 + conn_face(c)]) * a.tau_scale;
  }
  const double dt = a.mode == 0 ? *a.dtp : 0.0;
  double dtsum = 0.0;
  for (int v = 0; v < 5; v++) {
    for (int i = tid; i < NpA * EB; i += NT) s_r[v][i] = 0.0;  // padded rows and columns
  }
  __syncthreads();
#pragma unroll
  for (int v = 0; v < 5; v++)
    for (int i = tid; i < ne * Np; i += NT) {
      const size_t gi = v * KNp + (size_t)e0 * Np + i;
      const int e = i / Np, j = i - e * Np;
      const double uu = a.u[gi];
      s_u[v][i] = uu;
      double r;
      if (a.mode == 0) {
        const double ub = a.ubar[gi];
        const double kb = a.kbar[gi] + a.b * ub;
        a.kbar[gi] = a.a * kb;
        r = dt * kb;
        if (a.dtpart && dt != 0.0) {
          const double knext = (a.u_next[gi] - uu) / a.b;
          const double kcur = a.u_prev ? (uu - a.u_prev[gi]) / a.b_prev : 0.0;
          dtsum += (knext - a.a * kcur) / dt * kb;
        }
      } else {
        r = a.rbar[gi];
      }
      s_r[v][j * EB + e] = r;
    }
  if (a.mode == 0 && a.dtpart) {
    // deterministic block sum into this block's slot: a fixed shuffle tree in
    // each wave, then the wave sums in order
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) dtsum += __shfl_xor(dtsum, o, 64);
    if (lane == 0) red[w] = dtsum;
    __syncthreads();
    if (tid == 0) {
      double t = 0.0;
      for (int q = 0; q < NT / 64; q++) t += red[q];
      a.dtpart[blockIdx.x] += t;
    }
  }
  __syncthreads();

  // ---- phase 1: g = A^T rbar on MFMA, wave w = row tile w ----
  {
    double afr[KSA];
#pragma unroll
    for (int ks = 0; ks < KSA; ks++) afr[ks] = afragT[((size_t)w * KSA + ks) * 64 + lane];
    const int kk = lane >> 4, e = lane & 15;
#pragma unroll 1
    for (int v = 0; v < 5; v++) {
      typename P::Acc acc = {0, 0, 0, 0};
#pragma unroll
      for (int ks = 0; ks < KSA; ks++) acc = P::mfma(afr[ks], s_r[v][(4 * ks + kk) * EB + e], acc);
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const int k = 16 * w + 4 * r + kk;
        if (k < KD) s_g[v][k * EB + e] = acc[r];
      }
    }
  }
  __syncthreads();

  // ---- phase 2: transposed Jacobians per volume node ----
  for (int idx = tid; idx < ne * Np; idx += NT) {
    const int e = idx / Np, j = idx - e * Np;
    const double *m = &s_geo[e * GEO_STRIDE];
    double U[5], wv[5] = {0, 0, 0, 0, 0};
#pragma unroll
    for (int v = 0; v < 5; v++) U[v] = s_u[v][idx];
    // volume: sum_x J_x(u_j)^T h^x,  h^x_v = sum_d metric[d][x] g[d*Np+j][v]
#pragma unroll
    for (int x = 0; x < 3; x++) {
      double h[5];
#pragma unroll
      for (int v = 0; v < 5; v++)
        h[v] = m[0 + x] * s_g[v][j * EB + e] + m[3 + x] * s_g[v][(Np + j) * EB + e] +
               m[6 + x] * s_g[v][(2 * Np + j) * EB + e];
      euler_jt(U, x == 0 ? 1.0 : 0.0, x == 1 ? 1.0 : 0.0, x == 2 ? 1.0 : 0.0, h, a.gamma, wv);
      __builtin_amdgcn_sched_barrier(0);  // one direction at a time: bounds live registers
    }
    // own faces through this node (at most 3 slots)
#pragma unroll 1
    for (int q = 0; q < 3; q++) {
      const int mslot = s_slots[j * 3 + q];
      if (mslot < 0) break;
      const int f = mslot / Nfp, jj = mslot - f * Nfp;
      const int c = s_conn[e * 4 + f];
      const size_t pidx = (size_t)conn_nbr(c) * Np + s_fmx[(conn_face(c) * 4 + conn_code(c)) * Nfp + jj];
      const double *g = &s_geo[e * GEO_STRIDE + 9 + f * 4];
      const double tau = s_tau[e * 4 + f];
      double ub[5], h[5], ww[5] = {0, 0, 0, 0, 0};
#pragma unroll
      for (int v = 0; v < 5; v++) {
        ub[v] = 0.5 * (U[v] + a.u[v * KNp + pidx]);
        h[v] = 0.5 * g[3] * s_g[v][(3 * Np + mslot) * EB + e];
      }
      euler_jt(ub, g[0], g[1], g[2], h, a.gamma, ww);
#pragma unroll
      for (int v = 0; v < 5; v++) wv[v] += ww[v] + tau * h[v];
    }
#pragma unroll
    for (int v = 0; v < 5; v++) {
      const size_t gi = v * KNp + (size_t)e0 * Np + idx;
      a.ubar[gi] = (a.mode == 0 ? a.ubar[gi] : 0.0) + wv[v];
    }
  }
  // tau adjoint per face side: sum_m sum_v g[m][v] * (-(ub_v - u-_v) Fscale)
  if (a.taubar)
    for (int idx = tid; idx < ne * 4; idx += NT) {
      const int e = idx >> 2, f = idx & 3;
      const int c = s_conn[idx];
      const double fs = s_geo[e * GEO_STRIDE + 9 + f * 4 + 3];
      double t = 0.0;
#pragma unroll 1
      for (int j = 0; j < Nfp; j++) {
        const int mloc = e * Np + s_fmx[(f * 4 + 3) * Nfp + j];
        const size_t pidx = (size_t)conn_nbr(c) * Np + s_fmx[(conn_face(c) * 4 + conn_code(c)) * Nfp + j];
#pragma unroll
        for (int v = 0; v < 5; v++) {
          const double um = s_u[v][mloc];
          const double ub = 0.5 * (um + a.u[v * KNp + pidx]);
          t += s_g[v][(3 * Np + f * Nfp + j) * EB + e] * (-(ub - um) * fs);
        }
      }
      const size_t o = (size_t)e0 * 4 + idx;
      a.taubar[o] = (a.mode == 0 ? a.taubar[o] : 0.0) + t;
    }
  // face part of g for pass B, [K][NFF][5]
  for (int idx = tid; idx < ne * NFF * 5; idx += NT) {
    const int e = idx / (NFF * 5), r = idx - e * NFF * 5, mm = r / 5, v = r - mm * 5;
    a.gsur[(size_t)(e0 + e) * NFF * 5 + r] = s_g[v][(3 * Np + mm) * EB + e];
  }
}

}  // namespace dglm
