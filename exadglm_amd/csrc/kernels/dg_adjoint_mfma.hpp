// Pass A of the discrete Euler adjoint with the contraction g = A^T rbar on
// FP64 MFMA (gfx950).  Same mathematics as adjoint_pass_a (dg_adjoint.hpp),
// except that the dt partial sum_nodes kb . r is formed from the quantities
// of phase 2 (homogeneity of the Euler flux) rather than from differences of
// the stored stage states; the other difference is the layout of the work:
//
//   * a workgroup takes 16 elements (the MFMA columns) and has one wave per
//     16-row tile of A^T (KT = ceil(KD/16) waves, 11 at N=4); wave T keeps
//     the A^T fragments of its rows for the k-steps over the Np nodes in
//     registers and runs ceil(Np/4) v_mfma_f64_16x16x4f64 per variable;
//   * each thread owns one volume node of the 16 elements; phase 0 loads its
//     u, ubar and kbar into registers, adds to ubar the partner-side face
//     terms the previous stage's pass A left for the node (when given: the
//     work of the pass B of that stage), updates kbar and puts
//     rbar = dt (kbar + b ubar) into an LDS tile [node][element];
//   * phase 1 is the contraction, g[v][k][e] into LDS (all five variables,
//     since the transposed flux Jacobians mix them);
//   * phase 2 is pointwise per volume node: the transposed volume Jacobians,
//     the node's own-face terms (exterior traces gathered from global memory
//     ahead of the volume terms) and the write of ubar; each face-node slot
//     stores the partner-side term its neighbour's node needs (pass B) and
//     leaves its tau-bar partial in LDS, summed per face side afterwards.
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
  static constexpr int EBP = EB + 1;  // LDS row stride of the rbar and g tiles
};

template <int N>
constexpr size_t adj_mfma_lds_bytes() {
  using S = AdjShape<N>;
  return 8 * (5 * S::NpA * S::EBP + 5 * S::KD * S::EBP + S::EB * GEO_STRIDE + S::EB * 4 +
              S::NT / 64) +
         4 * (S::EB * 4 + 16 * S::Nfp + S::Np * 3 + 4 * S::Np);
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
  constexpr int EB = S::EB, EBP = S::EBP;
  static_assert(EB * Np <= NT, "one volume node per thread");
  // rows padded to EBP elements: the phase-2 reads (one node per lane, the
  // lanes of a wave on consecutive nodes) then fall in distinct banks
  __shared__ double s_r[5][NpA * EBP];  // rbar, [node][element]
  __shared__ double s_g[5][KD * EBP];   // A^T rbar, [row][element]
  __shared__ double s_geo[EB * GEO_STRIDE];
  __shared__ double s_tau[EB * 4];
  __shared__ int s_conn[EB * 4];
  __shared__ int s_fmx[16 * Nfp];
  __shared__ int s_slots[Np * 3];
  __shared__ int s_fpos[4 * Np];
  __shared__ double red[NT / 64];
  // tau-bar partial of each face-node slot, [element][slot]; rbar is dead
  // once the contraction is done
  static_assert(EB * NFF <= 5 * NpA * EBP, "slot partials must fit the rbar tile");
  double *const s_tc = &s_r[0][0];

  const int K = a.K;
  const size_t KNp = (size_t)K * Np;
  const int e0 = blockIdx.x * EB;
  const int ne = min(EB, K - e0);
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int w = __builtin_amdgcn_readfirstlane(tid >> 6);
  // thread tid owns volume node j of element e of the block for the whole
  // kernel: its state and ubar stay in registers from phase 0 to phase 2
  const bool live = tid < ne * Np;
  const int e = tid / Np, j = tid - e * Np;
  const size_t g0 = (size_t)e0 * Np + tid;

  // A^T fragments of the wave's row tile, first used in phase 1: issued
  // first so that their latency is hidden behind phase 0
  double afr[KSA];
#pragma unroll
  for (int ks = 0; ks < KSA; ks++) afr[ks] = afragT[((size_t)w * KSA + ks) * 64 + lane];

  // ---- phase 0: staging, rbar, kbar update ----
  for (int i = tid; i < 16 * Nfp; i += NT) s_fmx[i] = a.fmx[i];
  for (int i = tid; i < Np * 3; i += NT) s_slots[i] = a.node_slots[i];
  if (a.gsur_in)
    for (int i = tid; i < 4 * Np; i += NT) s_fpos[i] = a.fpos[i];
  for (int i = tid; i < ne * GEO_STRIDE; i += NT) s_geo[i] = a.geo[(size_t)e0 * GEO_STRIDE + i];
  // wave speeds for tau: loaded here, written to LDS after the contraction
  // (tau is first read in phase 2), so their dependent loads do not hold up
  // phase 0
  static_assert(EB * 4 <= 64, "one face side per lane of wave 0");
  double lam_own = 0.0, lam_nbr = 0.0;
  if (tid < ne * 4) {
    const int c = a.conn[(size_t)e0 * 4 + tid];
    s_conn[tid] = c;
    lam_own = a.maxv[(size_t)e0 * 4 + tid];
    lam_nbr = a.maxv[conn_nbr(c) * 4 + conn_face(c)];
  }
  for (int v = 0; v < 5; v++) {
    for (int i = tid; i < NpA * EBP; i += NT) s_r[v][i] = 0.0;  // padded rows and columns
  }
  const double dt = a.mode == 0 ? *a.dtp : 0.0;
  double U[5], UB[5], R[5];
  int cq[3] = {0, 0, 0};  // connectivity words of the node's faces (pending partner terms)
  if (live) {
    // every load before any store (kbar may alias nothing, but the compiler
    // cannot know): one round trip instead of five
#pragma unroll
    for (int v = 0; v < 5; v++) {
      U[v] = a.u[v * KNp + g0];
      if (a.mode == 0) {
        UB[v] = a.ubar[v * KNp + g0];
        R[v] = a.kbar[v * KNp + g0];
      } else {
        UB[v] = 0.0;
        R[v] = a.rbar[v * KNp + g0];
      }
    }
    if (a.gsur_in) {
#pragma unroll
      for (int q = 0; q < 3; q++) {
        const int ms = a.node_slots[j * 3 + q];
        if (ms >= 0) cq[q] = a.conn[(size_t)(e0 + e) * 4 + ms / Nfp];
      }
    }
  }
  __syncthreads();  // s_r zeroed; s_fmx, s_slots, s_fpos
  if (live && a.gsur_in) {
    // the previous stage's pass-B work: ubar += the partner-side terms the
    // neighbours' face-node slots stored for this node (same sum as
    // adjoint_pass_b, so ubar is bitwise what that pass would leave)
    double t[3][5];
#pragma unroll
    for (int q = 0; q < 3; q++) {
      const int ms = s_slots[j * 3 + q];
#pragma unroll
      for (int v = 0; v < 5; v++) t[q][v] = 0.0;
      if (ms >= 0) {
        const int f = ms / Nfp, jj = ms - f * Nfp;
        const int c = cq[q];
        const int nb = conn_nbr(c), nf = conn_face(c);
        const int pnode = s_fmx[(nf * 4 + conn_code(c)) * Nfp + jj];  // neighbour's volume node
        const int jp = s_fpos[nf * Np + pnode];                       // its slot on face nf
        const double *gs = &a.gsur_in[((size_t)nb * NFF + nf * Nfp + jp) * 5];
#pragma unroll
        for (int v = 0; v < 5; v++) t[q][v] = gs[v];
      }
    }
#pragma unroll
    for (int v = 0; v < 5; v++) {
      double wsum = 0.0;
#pragma unroll
      for (int q = 0; q < 3; q++) wsum += t[q][v];
      UB[v] += wsum;
    }
  }
  if (live && a.mode == 0) {
#pragma unroll
    for (int v = 0; v < 5; v++) {
      const double kb = R[v] + a.b * UB[v];
      a.kbar[v * KNp + g0] = a.a * kb;
      R[v] = dt * kb;
    }
  }
  if (live) {
#pragma unroll
    for (int v = 0; v < 5; v++) s_r[v][j * EBP + e] = R[v];
  }
  __syncthreads();

  // ---- phase 1: g = A^T rbar on MFMA, wave w = row tile w ----
  {
    const int kk = lane >> 4, ec = lane & 15;
#pragma unroll 1
    for (int v = 0; v < 5; v++) {
      typename P::Acc acc = {0, 0, 0, 0};
#pragma unroll
      for (int ks = 0; ks < KSA; ks++) acc = P::mfma(afr[ks], s_r[v][(4 * ks + kk) * EBP + ec], acc);
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const int k = 16 * w + 4 * r + kk;
        if (k < KD) s_g[v][k * EBP + ec] = acc[r];
      }
    }
  }
  if (tid < ne * 4) s_tau[tid] = 0.25 * (lam_own + lam_nbr) * a.tau_scale;
  __syncthreads();

  // ---- phase 2: transposed Jacobians at the thread's volume node ----
  // With the dt partial: sum over rows of g . B(u) = sum_nodes rbar . r, and
  // B is known from what is formed here because the Euler flux is
  // homogeneous of degree one, F(U) = J(U) U: the volume part is
  // sum_x h^x . F^x(U) = (sum_x J_x^T h^x) . U, the face part of slot m is
  // g_m . (n.F(ub) - tau (ub - u-)) Fscale = 2 (J_n(ub)^T h_m) . ub + tau tc_m.
  double rdot = 0.0;
  if (live) {
    const double *m = &s_geo[e * GEO_STRIDE];
    // exterior traces of the (at most 3) faces through this node, gathered
    // first so that their latency overlaps the volume terms
    int ms[3];
    double uP[3][5];
#pragma unroll
    for (int q = 0; q < 3; q++) {
      ms[q] = s_slots[j * 3 + q];
      if (ms[q] >= 0) {
        const int f = ms[q] / Nfp, jj = ms[q] - f * Nfp;
        const int c = s_conn[e * 4 + f];
        const size_t pidx = (size_t)conn_nbr(c) * Np + s_fmx[(conn_face(c) * 4 + conn_code(c)) * Nfp + jj];
#pragma unroll
        for (int v = 0; v < 5; v++) uP[q][v] = a.u[v * KNp + pidx];
      }
    }
    double wv[5] = {0, 0, 0, 0, 0};
    // volume: sum_x J_x(u_j)^T h^x,  h^x_v = sum_d metric[d][x] g[d*Np+j][v]
#pragma unroll
    for (int x = 0; x < 3; x++) {
      double h[5];
#pragma unroll
      for (int v = 0; v < 5; v++)
        h[v] = m[0 + x] * s_g[v][j * EBP + e] + m[3 + x] * s_g[v][(Np + j) * EBP + e] +
               m[6 + x] * s_g[v][(2 * Np + j) * EBP + e];
      euler_jt(U, x == 0 ? 1.0 : 0.0, x == 1 ? 1.0 : 0.0, x == 2 ? 1.0 : 0.0, h, a.gamma, wv);
      __builtin_amdgcn_sched_barrier(0);  // one direction at a time: bounds live registers
    }
#pragma unroll
    for (int v = 0; v < 5; v++) rdot += wv[v] * U[v];
    // own faces through this node; the face-node slot's tau-bar partial
    // tc = sum_v g[m][v] * (-(ub_v - u-_v) Fscale) goes to LDS
#pragma unroll
    for (int q = 0; q < 3; q++) {
      if (ms[q] < 0) break;
      const int mslot = ms[q];
      const int f = mslot / Nfp;
      const double *g = &s_geo[e * GEO_STRIDE + 9 + f * 4];
      const double tau = s_tau[e * 4 + f];
      double ub[5], h[5], ww[5] = {0, 0, 0, 0, 0};
      double tc = 0.0;
#pragma unroll
      for (int v = 0; v < 5; v++) {
        const double gm = s_g[v][(3 * Np + mslot) * EBP + e];
        ub[v] = 0.5 * (U[v] + uP[q][v]);
        h[v] = 0.5 * g[3] * gm;
        tc += gm * (-(ub[v] - U[v]) * g[3]);
      }
      s_tc[e * NFF + mslot] = tc;
      euler_jt(ub, g[0], g[1], g[2], h, a.gamma, ww);
      double fw = 0.0;
      // the partner's node gets (1/2 n.J - tau/2)^T Fscale g: stored for pass B
      double *pt = &a.gsur[((size_t)(e0 + e) * NFF + mslot) * 5];
#pragma unroll
      for (int v = 0; v < 5; v++) {
        wv[v] += ww[v] + tau * h[v];
        pt[v] = ww[v] - tau * h[v];
        fw += ww[v] * ub[v];
      }
      rdot += 2.0 * fw + tau * tc;
      __builtin_amdgcn_sched_barrier(0);
    }
#pragma unroll
    for (int v = 0; v < 5; v++) a.ubar[v * KNp + g0] = UB[v] + wv[v];
  }
  if (a.mode == 0 && a.dtpart && dt != 0.0) {
    // deterministic block sum of rdot/dt = sum_nodes kb . r into this block's
    // slot: a fixed shuffle tree in each wave, then the wave sums in order
    double dtsum = rdot / dt;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) dtsum += __shfl_xor(dtsum, o, 64);
    if (lane == 0) red[w] = dtsum;
  }
  // tau adjoint per face side: the slot partials in face-node order
  if (a.taubar || (a.mode == 0 && a.dtpart && dt != 0.0)) __syncthreads();
  if (a.mode == 0 && a.dtpart && dt != 0.0 && tid == 0) {
    double t = 0.0;
    for (int q = 0; q < NT / 64; q++) t += red[q];
    a.dtpart[blockIdx.x] += t;
  }
  if (a.taubar)
    for (int idx = tid; idx < ne * 4; idx += NT) {
      const int el = idx >> 2, f = idx & 3;
      double t = 0.0;
#pragma unroll
      for (int jj = 0; jj < Nfp; jj++) t += s_tc[el * NFF + f * Nfp + jj];
      const size_t o = (size_t)e0 * 4 + idx;
      a.taubar[o] = (a.mode == 0 ? a.taubar[o] : 0.0) + t;
    }
}

}  // namespace dglm
