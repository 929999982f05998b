// FP64 MFMA stage kernel (gfx950): persistent workgroups, operator resident
// in registers, per-variable flux tiles in LDS, LSRK update as the epilogue.
//
// The stage contraction  rhs[v][e][i] = sum_k A[i][k] B_v[k][e]  with
// A = [WDr | WDs | WDt | -LIFT] (Np x KD) and B_v the stacked volume fluxes
// (rows d*Np + j) and LM surface fluxes (rows 3Np + m) of variable v is a
// tiny-M / huge-N product (update_rhs, dim3/src/cpp/euler_dglm.cpp:254-302;
// 4 cublasDgemmBatched calls in euler_nopad_dglm.cu:269-300).  Here it runs
// on v_mfma_f64_16x16x4f64:
//   * one workgroup = RT = ceil(Np/16) waves; wave w owns rows 16w..16w+15 and
//     keeps its KS = ceil(KD/4) A-fragments in registers for the whole kernel;
//   * a batch = 16 elements = the 16 MFMA columns; for each variable the
//     workgroup writes B_v into an LDS tile [KDP][16] (k-major, element
//     fastest: a B-fragment read is 64 consecutive doubles, conflict free);
//   * the D-fragment (register r of lane l holds row 4r + l/16, column l%16;
//     the f64 16x16x4 layout, unlike the f32 one) is consumed in
//     registers by the LSRK update k = a k + dt rhs, u_out = u_in + b k.
// Workgroups loop over batches (persistent), so the operator is read once
// per workgroup.
#pragma once
#include "dg_fp64.hpp"

namespace dglm {

typedef double f64x4 __attribute__((ext_vector_type(4)));

template <int N>
struct MfmaShape {
  using O = Ord<N>;
  static constexpr int RT = (O::Np + 15) / 16;  // row tiles = waves per workgroup
  static constexpr int KS = (O::KD + 3) / 4;    // k-steps of 4
  static constexpr int KDP = KS * 4;            // padded contraction length
  static constexpr int NT = RT * 64;            // threads per workgroup
};

// Host-side arrangement of the stacked operator into A-fragments:
// frag[(w*KS + ks)*64 + lane] = A[16w + lane%16][4ks + lane/16] (0 outside).
inline void arrange_a_fragments(int Np, int KD, int RT, int KS, const double *opT /*[KD][Np]*/, double *frag) {
  for (int w = 0; w < RT; w++)
    for (int ks = 0; ks < KS; ks++)
      for (int l = 0; l < 64; l++) {
        int i = 16 * w + (l & 15), k = 4 * ks + (l >> 4);
        frag[((size_t)w * KS + ks) * 64 + l] = (i < Np && k < KD) ? opT[(size_t)k * Np + i] : 0.0;
      }
}

//   MODE 0: LSRK update; MODE 1: write the RHS only.
template <int N, int PHYS, int MODE>
__global__ __launch_bounds__(MfmaShape<N>::NT) void stage_mfma_kernel(StageArgs a, const double *__restrict__ afrag,
                                                                      int nbatch) {
  using O = Ord<N>;
  using S = MfmaShape<N>;
  constexpr int Np = O::Np, Nfp = O::Nfp, NFF = O::NFF, KD = O::KD;
  constexpr int RT = S::RT, KS = S::KS, KDP = S::KDP, NT = S::NT;
  constexpr int NV = NVar<PHYS>::value;
  constexpr int EB = 16;
  constexpr int NVEL = PHYS == PHYS_ADV ? 3 : 1;

  __shared__ double s_u[NV][EB * Np];
  __shared__ double s_f[KDP * EB];
  __shared__ double s_vel[NVEL][PHYS == PHYS_ADV ? EB * Np : 1];
  __shared__ double s_geo[EB * GEO_STRIDE];
  __shared__ double s_tau[EB * 4];
  __shared__ int s_conn[EB * 4];
  __shared__ int s_fmx[16 * Nfp];

  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int w = tid >> 6;
  const int K = a.K;
  const size_t KNp = (size_t)K * Np;

  if (MODE == 0 && !a.scal->active) {
    // Past tmax: the stage is the identity on u (k untouched).
    for (size_t i = (size_t)blockIdx.x * NT + tid; i < NV * KNp; i += (size_t)gridDim.x * NT) a.u_out[i] = a.u_in[i];
    return;
  }
  const double dt = MODE == 0 ? a.scal->dt : 0.0;

  // Operator rows of this wave, resident for the whole kernel.
  double afr[KS];
#pragma unroll
  for (int ks = 0; ks < KS; ks++) afr[ks] = afrag[((size_t)w * KS + ks) * 64 + lane];

  for (int i = tid; i < 16 * Nfp; i += NT) s_fmx[i] = a.fmx[i];
  for (int i = KD * EB + tid; i < KDP * EB; i += NT) s_f[i] = 0.0;  // padded rows stay zero

  for (int b = blockIdx.x; b < nbatch; b += gridDim.x) {
    const int e0 = b * EB;
    const int ne = min(EB, K - e0);
    __syncthreads();  // previous batch finished with s_u / s_geo
    for (int i = tid; i < ne * GEO_STRIDE; i += NT) s_geo[i] = a.geo[(size_t)e0 * GEO_STRIDE + i];
    for (int i = tid; i < ne * 4; i += NT) {
      const int c = a.conn[(size_t)e0 * 4 + i];
      s_conn[i] = c;
      s_tau[i] = 0.25 * (a.maxv[(size_t)e0 * 4 + i] + a.maxv[conn_nbr(c) * 4 + conn_face(c)]) * a.tau_scale;
    }
#pragma unroll
    for (int v = 0; v < NV; v++)
      for (int i = tid; i < ne * Np; i += NT) s_u[v][i] = a.u_in[v * KNp + (size_t)e0 * Np + i];
    if (PHYS == PHYS_ADV)
      for (int d = 0; d < 3; d++)
        for (int i = tid; i < ne * Np; i += NT) s_vel[d][i] = a.vel[d * KNp + (size_t)e0 * Np + i];
    __syncthreads();

#pragma unroll 1
    for (int v = 0; v < NV; v++) {
      // ---- B_v tile: volume rows (d*Np + j), column e ----
      for (int idx = tid; idx < Np * EB; idx += NT) {
        const int j = idx >> 4, e = idx & 15;
        double f0 = 0.0, f1 = 0.0, f2 = 0.0;
        if (e < ne) {
          const double *g = &s_geo[e * GEO_STRIDE];
          double F, G, H;
          if (PHYS == PHYS_EULER) {
            const double u1 = s_u[0][e * Np + j], u2 = s_u[1][e * Np + j], u3 = s_u[2][e * Np + j],
                         u4 = s_u[3][e * Np + j], u5 = s_u[4][e * Np + j];
            const double inv = 1.0 / u1;
            const double p = (a.gamma - 1.0) * (u5 - 0.5 * (u2 * u2 + u3 * u3 + u4 * u4) * inv);
            const double vx = u2 * inv, vy = u3 * inv, vz = u4 * inv;
            if (v == 0) {
              F = u2;
              G = u3;
              H = u4;
            } else {
              // momentum component (v = 1..3) or total enthalpy density (v = 4)
              // times velocity, plus the pressure on the diagonal
              const double q = v == 4 ? u5 + p : (v == 1 ? u2 : v == 2 ? u3 : u4);
              F = q * vx + (v == 1 ? p : 0.0);
              G = q * vy + (v == 2 ? p : 0.0);
              H = q * vz + (v == 3 ? p : 0.0);
            }
          } else {
            const double uu = s_u[0][e * Np + j];
            F = s_vel[0][e * Np + j] * uu;
            G = s_vel[1][e * Np + j] * uu;
            H = s_vel[2][e * Np + j] * uu;
          }
          f0 = g[0] * F + g[1] * G + g[2] * H;
          f1 = g[3] * F + g[4] * G + g[5] * H;
          f2 = g[6] * F + g[7] * G + g[8] * H;
        }
        s_f[j * EB + e] = f0;
        s_f[(Np + j) * EB + e] = f1;
        s_f[(2 * Np + j) * EB + e] = f2;
      }
      // ---- B_v tile: LM surface rows (3Np + m) ----
      for (int idx = tid; idx < NFF * EB; idx += NT) {
        const int m = idx >> 4, e = idx & 15;
        double fl = 0.0;
        if (e < ne) {
          const int f = m / Nfp, j = m - f * Nfp;
          const int c = s_conn[e * 4 + f];
          const size_t pidx = (size_t)conn_nbr(c) * Np + s_fmx[(conn_face(c) * 4 + conn_code(c)) * Nfp + j];
          const int mloc = e * Np + s_fmx[(f * 4 + 3) * Nfp + j];
          const double *g = &s_geo[e * GEO_STRIDE + 9 + f * 4];
          const double tau = s_tau[e * 4 + f];
          if (PHYS == PHYS_EULER) {
            double ub[5];
#pragma unroll
            for (int q = 0; q < 5; q++) ub[q] = 0.5 * (s_u[q][mloc] + a.u_in[q * KNp + pidx]);
            const double inv = 1.0 / ub[0];
            const double p = (a.gamma - 1.0) * (ub[4] - 0.5 * (ub[1] * ub[1] + ub[2] * ub[2] + ub[3] * ub[3]) * inv);
            const double vn = (g[0] * ub[1] + g[1] * ub[2] + g[2] * ub[3]) * inv;  // n . velocity
            double nF;
            if (v == 0) nF = g[0] * ub[1] + g[1] * ub[2] + g[2] * ub[3];
            else if (v == 4) nF = (ub[4] + p) * vn;
            else nF = ub[v] * vn + g[v - 1] * p;
            fl = (nF - tau * (ub[v] - s_u[v][mloc])) * g[3];
          } else {
            const double uM = s_u[0][mloc];
            const double ub = 0.5 * (uM + a.u_in[pidx]);
            const double nF = g[0] * (s_vel[0][mloc] * ub) + g[1] * (s_vel[1][mloc] * ub) + g[2] * (s_vel[2][mloc] * ub);
            fl = (nF - tau * (ub - uM)) * g[3];
          }
        }
        s_f[(3 * Np + m) * EB + e] = fl;
      }
      __syncthreads();

      // ---- contraction on FP64 MFMA ----
      f64x4 acc = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
      for (int ks = 0; ks < KS; ks++) {
        const double bf = s_f[(ks * 4 + (lane >> 4)) * EB + (lane & 15)];
        acc = __builtin_amdgcn_mfma_f64_16x16x4f64(afr[ks], bf, acc, 0, 0, 0);
      }

      // ---- LSRK epilogue: rows 16w + 4r + lane/16, element lane%16 ----
      const int e = lane & 15;
      if (e < ne) {
#pragma unroll
        for (int r = 0; r < 4; r++) {
          const int i = 16 * w + 4 * r + (lane >> 4);
          if (i < Np) {
            const size_t gi = v * KNp + (size_t)(e0 + e) * Np + i;
            if (MODE == 1) {
              a.rhs[gi] = acc[r];
            } else {
              const double kn = a.rka * a.k[gi] + dt * acc[r];
              a.k[gi] = kn;
              a.u_out[gi] = s_u[v][e * Np + i] + a.rkb * kn;
            }
          }
        }
      }
      __syncthreads();  // s_f is rewritten for the next variable
    }
  }
}

}  // namespace dglm
