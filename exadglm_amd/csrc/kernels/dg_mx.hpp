// Mixed-precision Euler stage kernel (gfx950): the operator contraction on
// FP32 MFMA, applied to what a per-block least-squares affine fit leaves of
// each flux block; the fit is contracted exactly in FP64.
//
// The stage contraction  rhs[v][e][i] = sum_k A[i][k] B_v[k][e]  (update_rhs,
// dim3/src/cpp/euler_dglm.cpp:254-302; A = [WDr|WDs|WDt|-LIFT], B_v the
// stacked volume fluxes and surface fluxes of variable v) of a smooth field
// is a small difference of large volume and surface terms, so a plain FP32
// contraction loses ~1e-7 relative.  Each flux block b (the three volume
// directions over the element's nodes, each face over its nodes) is written
// as an affine function V_b c_b plus a remainder:
//   B_b = V_b c_b + R_b,
//   rhs = A R (FP32 MFMA, FP32 accumulation) + (A V) c (FP64 MFMA),
// with c_b = P_b^T B_b the least-squares fit (FP64; V, P from the reference
// nodes, the tables of affine_fit_tables).  The identity holds for any c, so
// the split changes rounding only: only R is rounded to FP32.  The LSRK state
// and update, pressure, wave speed, fluxes, fits and the (A V) c term stay
// FP64.
//
// Organisation: one persistent workgroup per CU of eight waves (wave w runs
// on SIMD w % 4): RT = ceil(Np/16) contraction waves and 8 - RT helpers.
// Contraction wave t keeps the FP32 A-fragments of rows 16t..16t+15 (all
// k-steps) and its rows of (A V) in registers for the whole kernel.  Per
// 16-element batch (the 16 MFMA columns):
//   1. the helpers write the batch's node values, geometry and tau into LDS
//      (they loaded them into registers during the previous batch's variable
//      loop);
//   2. the surface flux of all five variables at every face-node pair, once,
//      its face fits, and the remainders rounded into an FP32 tile per
//      variable (eight threads per face; waves 0..3 take faces 0 and 1,
//      waves 4..7 faces 2 and 3, whose exterior traces lie farther away);
//   3. variable 0's volume pass (volume fluxes, their fits, the remainder
//      tile), on waves 0..3, which finish step 2 first;
//   4. per variable v: the contraction waves run the MFMA contraction over
//      [volume | surface] k-steps, the fit term (A V) c on FP64 MFMA and the
//      LSRK update, while the volume pass of variable v + 1 goes into the
//      other remainder buffer; one barrier per variable.
// A volume pass is eight slots of two elements each.  In step 4 wave w takes
// slot w, except at RT = 6 (N = 6), where contraction waves 4 and 5 share
// SIMDs 0 and 1 with waves 0 and 1 and their slots go to the helpers.
//
// No Dirichlet faces and no advection here: those take the generic MFMA
// kernel (dg_mfma.hpp).
#pragma once
#include <vector>

#include "dg_mfma.hpp"

namespace dglm {

template <int N>
struct MxShape {
  using O = Ord<N>;
  static constexpr int Np = O::Np, Nfp = O::Nfp, NFF = O::NFF, KD = O::KD;
  static constexpr int RT = (Np + 15) / 16;     // row tiles = contraction waves
  static constexpr int NW = 8;                  // waves: RT contraction waves + 8 - RT helpers
  static constexpr int NT = 64 * NW;
  static constexpr int KSV = (3 * Np + 3) / 4;  // volume k-steps (rows 3Np..4KSV-1 are zero)
  static constexpr int KV4 = 4 * KSV;
  static constexpr int KS = KSV + Nfp;  // + surface k-steps (NFF = 4 Nfp rows)
  static constexpr int EB = 16;
  static constexpr int RS = EB + 1;  // row stride of the FP32 volume remainder tile (odd: a volume pass
                                     // writes one element at 32 consecutive rows, which then spread over
                                     // the banks); the surface tile keeps EB (its reads are then
                                     // conflict-free and its writes, 8 elements at 8 face rows, 2-way)
  static constexpr int NC = 24;  // fit coefficients per (variable, element): 3 x 4 volume + 4 x 3 face
  static constexpr int UST = Np + 1;  // element stride of the staged node values (odd: a 16-element
                                      // column spreads over the LDS banks)
};

// Device tables of the kernel (host-built, one buffer):
//   afr  [RT][KS][64] float    A-fragments, k order [volume rows padded to KV4 | surface rows]
//   avf  [RT][NC/4][64] double FP64 A-fragments of (A V), row i's weight of each
//                              fit coefficient, rows permuted so that the FP64
//                              MFMA's D layout (row 4r + l/16) matches the FP32
//                              one (row 4(l/16) + r): fragment row 4r + q holds
//                              row 4q + r of the tile
//   pv, vv [Np][4] double      volume fit weights and basis (the three
//                              directions share the node set)
//   pf, vf [NFF][4] double     face fit weights and basis (q = 3 unused)
template <int N>
struct MxTables {
  using S = MxShape<N>;
  static constexpr size_t afr_off = 0;
  static constexpr size_t afr_bytes = sizeof(float) * S::RT * S::KS * 64;
  static constexpr size_t avf_off = (afr_bytes + 255) / 256 * 256;
  static constexpr size_t pv_off = avf_off + sizeof(double) * S::RT * S::NC * 16;
  static constexpr size_t vv_off = pv_off + sizeof(double) * S::Np * 4;
  static constexpr size_t pf_off = vv_off + sizeof(double) * S::Np * 4;
  static constexpr size_t vf_off = pf_off + sizeof(double) * S::NFF * 4;
  static constexpr size_t bytes = vf_off + sizeof(double) * S::NFF * 4;
};

// Host-side: fill the tables of order N from the stacked operator opT
// [KD][Np] (op[k][i] = A[i][k], -LIFT included) and the affine-fit tables
// fitV, fitP [KD][4], fitAV [Np][NBLK][4] of affine_fit_tables.
template <int N>
inline void mx_build_tables(const double *opT, const double *fitV, const double *fitP, const double *fitAV,
                            unsigned char *buf) {
  using S = MxShape<N>;
  constexpr int Np = S::Np;
  using T = MxTables<N>;
  float *afr = (float *)(buf + T::afr_off);
  double *avf = (double *)(buf + T::avf_off), *pv = (double *)(buf + T::pv_off), *vv = (double *)(buf + T::vv_off);
  double *pf = (double *)(buf + T::pf_off), *vf = (double *)(buf + T::vf_off);
  for (int j = 0; j < Np; j++)
    for (int q = 0; q < 4; q++) {
      pv[j * 4 + q] = fitP[j * 4 + q];  // block 0 (rows 0..Np-1); blocks 1, 2 are the same nodes
      vv[j * 4 + q] = fitV[j * 4 + q];
    }
  for (int m = 0; m < S::NFF; m++)
    for (int q = 0; q < 4; q++) {
      pf[m * 4 + q] = fitP[(3 * Np + m) * 4 + q];
      vf[m * 4 + q] = fitV[(3 * Np + m) * 4 + q];
    }
  // A-fragments in the kernel's k order
  auto kmap = [&](int kk) -> int {  // kernel row -> stacked row (-1: zero)
    if (kk < S::KV4) return kk < 3 * Np ? kk : -1;
    const int m = kk - S::KV4;
    return m < S::NFF ? 3 * Np + m : -1;
  };
  for (int w = 0; w < S::RT; w++)
    for (int ks = 0; ks < S::KS; ks++)
      for (int l = 0; l < 64; l++) {
        const int i = 16 * w + (l & 15), k = kmap(4 * ks + (l >> 4));
        afr[((size_t)w * S::KS + ks) * 64 + l] = (i < Np && k >= 0) ? (float)opT[(size_t)k * Np + i] : 0.0f;
      }
  // (A V): coefficient q' = d*4 + q (volume direction d) or 12 + f*3 + q (face f)
  auto av = [&](int i, int qq) {
    if (qq < 12) return fitAV[((size_t)i * NBLK + qq / 4) * 4 + qq % 4];
    const int f = (qq - 12) / 3, q = (qq - 12) % 3;
    return fitAV[((size_t)i * NBLK + 3 + f) * 4 + q];
  };
  for (int w = 0; w < S::RT; w++)
    for (int ks = 0; ks < S::NC / 4; ks++)
      for (int l = 0; l < 64; l++) {
        const int rho = l & 15, i = 16 * w + 4 * (rho & 3) + (rho >> 2), qq = 4 * ks + (l >> 4);
        avf[((size_t)w * (S::NC / 4) + ks) * 64 + l] = i < Np ? av(i, qq) : 0.0;
      }
}

// Two doubles at an 8-byte aligned address: one 16-byte memory instruction.
typedef double f64x2u __attribute__((ext_vector_type(2), aligned(8)));

// Workgroup barrier over LDS: waits for this wave's LDS (and scalar memory)
// operations only, not for its global loads and stores, so those stay in
// flight across it (no wave of the workgroup reads global memory another wave
// writes during the launch: the stage reads u_in and writes k and u_out).  A
// plain __syncthreads would also drain every outstanding global access.
__device__ inline void mx_barrier() { asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory"); }

// Surface flux of the five variables times Fscale at one face-node pair:
// LM (the reference's, euler_dglm.cpp:110-183) or Rusanov.  g = (nx, ny, nz,
// Fscale) of the face.
// euler_normal_flux5 / euler_rusanov5 (dg_fp64.hpp) with rcp_refined and
// sqrt_refined: the surface and volume passes are the critical paths of the
// stage-in and the variable loop
__device__ inline void mx_normal_flux5(const double U[5], const double *n, double gamma, double F[5], double &lam) {
  const double inv = rcp_refined(U[0]);
  const double p = (gamma - 1.0) * (U[4] - 0.5 * (U[1] * U[1] + U[2] * U[2] + U[3] * U[3]) * inv);
  const double vn = (n[0] * U[1] + n[1] * U[2] + n[2] * U[3]) * inv;
  lam = fabs(vn) + sqrt_refined(fabs(gamma * p * inv));
  F[0] = vn * U[0];
  F[1] = U[1] * vn + n[0] * p;
  F[2] = U[2] * vn + n[1] * p;
  F[3] = U[3] * vn + n[2] * p;
  F[4] = (U[4] + p) * vn;
}

__device__ inline void mx_surface_flux5(const double uM[5], const double uP[5], const double *g, double tau,
                                        int flux, double gamma, double fl[5]) {
  if (flux == 1) {
    double fm[5], fp[5], lm, lp;
    mx_normal_flux5(uM, g, gamma, fm, lm);
    mx_normal_flux5(uP, g, gamma, fp, lp);
    const double lam = fmax(lm, lp);
#pragma unroll
    for (int v = 0; v < 5; v++) fl[v] = 0.5 * (fm[v] + fp[v]) + 0.5 * lam * (uM[v] - uP[v]);
#pragma unroll
    for (int v = 0; v < 5; v++) fl[v] *= g[3];
    return;
  }
  double ub[5];
#pragma unroll
  for (int v = 0; v < 5; v++) ub[v] = 0.5 * (uM[v] + uP[v]);
  const double inv = rcp_refined(ub[0]);
  const double p = (gamma - 1.0) * (ub[4] - 0.5 * (ub[1] * ub[1] + ub[2] * ub[2] + ub[3] * ub[3]) * inv);
  const double vn = (g[0] * ub[1] + g[1] * ub[2] + g[2] * ub[3]) * inv;
  double nF[5];
  nF[0] = g[0] * ub[1] + g[1] * ub[2] + g[2] * ub[3];
  nF[1] = ub[1] * vn + g[0] * p;
  nF[2] = ub[2] * vn + g[1] * p;
  nF[3] = ub[3] * vn + g[2] * p;
  nF[4] = (ub[4] + p) * vn;
#pragma unroll
  for (int v = 0; v < 5; v++) fl[v] = (nF[v] - tau * (ub[v] - uM[v])) * g[3];
}

// Cross-lane sums without the LDS crossbar: a double moved by one DPP
// permutation (two 32-bit moves), and all-reduce sums over aligned groups of
// 8 lanes (quad swaps, then the half-row mirror) and of 32 lanes (rotations
// within a 16-lane row, then a swizzle with xor 16).  Every lane of a group
// ends with the same sum, formed in a fixed order (deterministic).
template <int CTRL>
__device__ inline double dpp_f64(double x) {
  const unsigned long long u = __builtin_bit_cast(unsigned long long, x);
  const unsigned lo = __builtin_amdgcn_update_dpp(0u, (unsigned)u, CTRL, 0xF, 0xF, false);
  const unsigned hi = __builtin_amdgcn_update_dpp(0u, (unsigned)(u >> 32), CTRL, 0xF, 0xF, false);
  return __builtin_bit_cast(double, ((unsigned long long)hi << 32) | lo);
}
__device__ inline double swz_xor16_f64(double x) {
  const unsigned long long u = __builtin_bit_cast(unsigned long long, x);
  constexpr int PAT = (0x10 << 10) | 0x1F;  // bit mode: and 0x1f, or 0, xor 0x10
  const unsigned lo = (unsigned)__builtin_amdgcn_ds_swizzle((int)(unsigned)u, PAT);
  const unsigned hi = (unsigned)__builtin_amdgcn_ds_swizzle((int)(unsigned)(u >> 32), PAT);
  return __builtin_bit_cast(double, ((unsigned long long)hi << 32) | lo);
}
__device__ inline double sum8(double x) {
  x += dpp_f64<0xB1>(x);   // quad_perm [1,0,3,2]
  x += dpp_f64<0x4E>(x);   // quad_perm [2,3,0,1]
  return x + dpp_f64<0x141>(x);  // row_half_mirror: the other quad of the 8
}
__device__ inline double sum32(double x) {
  x += dpp_f64<0x128>(x);  // row_ror 8
  x += dpp_f64<0x124>(x);  // row_ror 4
  x += dpp_f64<0x122>(x);  // row_ror 2
  x += dpp_f64<0x121>(x);  // row_ror 1
  return x + swz_xor16_f64(x);
}

// Volume fluxes (fR, fS, fT) of variable v at one node: g_d . F with the
// contravariant velocity V_d = g_d . v (no array indexed by v, so nothing
// here goes through private memory).
__device__ inline void mx_volume_flux1(int v, const double u[5], const double *gm, double gamma, double f[3]) {
  const double r = u[0], mx = u[1], my = u[2], mz = u[3], E = u[4];
  const double inv = rcp_refined(r);
  const double p = (gamma - 1.0) * (E - 0.5 * (mx * mx + my * my + mz * mz) * inv);
  const double vx = mx * inv, vy = my * inv, vz = mz * inv;
  const double mv = v == 1 ? mx : (v == 2 ? my : mz);  // momentum component of v = 1..3
#pragma unroll
  for (int d = 0; d < 3; d++) {
    const double g0 = gm[3 * d], g1 = gm[3 * d + 1], g2 = gm[3 * d + 2];
    const double V = g0 * vx + g1 * vy + g2 * vz;
    const double gv = v == 1 ? g0 : (v == 2 ? g1 : g2);
    if (v == 0)
      f[d] = g0 * mx + g1 * my + g2 * mz;
    else if (v == 4)
      f[d] = (E + p) * V;
    else
      f[d] = mv * V + gv * p;
  }
}

// Volume pass of wave w for variable v: the volume fluxes at every node of
// elements 2w, 2w+1 of the batch, their least-squares fits c (FP64, into cv)
// and the FP32 remainders into vr [KV4][EB].  Lane (element el, group sl)
// takes nodes j = sl + 32x; the fit sums meet across the 32 groups in
// registers.  pv, vv: fit weights and basis as [q][Np] (a read of one q by
// 32 consecutive nodes spreads over the LDS banks; [Np][4] put them on 8).
template <int N>
__device__ inline void mx_volume_pass(int v, int w, int lane, int ne, const double *s_u, const double *s_geo,
                                      const double *pv, const double *vv, double gamma, double *cv, float *vr) {
  using S = MxShape<N>;
  constexpr int Np = S::Np, EB = S::EB, UST = S::UST, JX = (Np + 31) / 32;
  const int el = 2 * w + (lane >> 5), sl = lane & 31;
  const double *gm = &s_geo[el * GEO_STRIDE];
  double f[JX][3];
  double c[12];
#pragma unroll
  for (int q = 0; q < 12; q++) c[q] = 0.0;
#pragma unroll
  for (int x = 0; x < JX; x++) {
    const int j = sl + 32 * x;
    f[x][0] = f[x][1] = f[x][2] = 0.0;
    if (j < Np) {
      double u[5];
#pragma unroll
      for (int q = 0; q < 5; q++) u[q] = s_u[q * EB * UST + el * UST + j];
      mx_volume_flux1(v, u, gm, gamma, f[x]);
#pragma unroll
      for (int d = 0; d < 3; d++)
#pragma unroll
        for (int q = 0; q < 4; q++) c[d * 4 + q] += pv[q * Np + j] * f[x][d];
    }
  }
#pragma unroll
  for (int q = 0; q < 12; q++) c[q] = sum32(c[q]);
  if (sl == 0)
#pragma unroll
    for (int q = 0; q < 12; q++) cv[q * EB + el] = c[q];
#pragma unroll
  for (int x = 0; x < JX; x++) {
    const int j = sl + 32 * x;
    if (j < Np) {
      const double V[4] = {vv[j], vv[Np + j], vv[2 * Np + j], vv[3 * Np + j]};
#pragma unroll
      for (int d = 0; d < 3; d++) {
        const double fit = V[0] * c[d * 4] + V[1] * c[d * 4 + 1] + V[2] * c[d * 4 + 2] + V[3] * c[d * 4 + 3];
        vr[(d * Np + j) * S::RS + el] = el < ne ? (float)(f[x][d] - fit) : 0.0f;
      }
    }
  }
}

//   MODE 0: LSRK update; MODE 1: write the RHS only.
// order: batch visiting order (null: mesh order); the grid is split into ngrp
// groups (blockIdx % ngrp) that walk contiguous stretches of it.
template <int N, int MODE>
__global__ __launch_bounds__(MxShape<N>::NT) void stage_mx_kernel(StageArgs a, const unsigned char *__restrict__ tab,
                                                                  const int *__restrict__ order, int nbatch,
                                                                  int ngrp) {
  using S = MxShape<N>;
  using TB = MxTables<N>;
  constexpr int Np = S::Np, Nfp = S::Nfp, NFF = S::NFF, KS = S::KS, KSV = S::KSV, KV4 = S::KV4;
  constexpr int NT = S::NT, RT = S::RT, EB = S::EB, NC = S::NC, UST = S::UST;
  constexpr int JF = (Nfp + 7) / 8;              // face nodes per thread in the surface pass
  static_assert(S::NW * 2 == EB && RT <= S::NW, "a volume pass takes two elements per wave");
  static_assert(NT == 4 * EB * 8, "the surface pass takes a thread per (face, group of eight)");
  static_assert(S::NW > RT && EB * 4 <= 64, "a helper wave stages the connectivity");
  const float *afrag = (const float *)(tab + TB::afr_off);
  const double *avfrag = (const double *)(tab + TB::avf_off);

  __shared__ double s_u[5][EB * UST];     // node values of the batch [v][e*UST + j]
  __shared__ float s_vr[2][KV4 * S::RS];  // volume remainders of two variables [k][e]
  __shared__ float s_sr[5][NFF * EB];     // surface remainders of all variables [v][m][e]
  __shared__ double s_c[5][NC * EB];      // fit coefficients [v][q'][e]
  __shared__ double s_tabv[2][Np * 4];    // fit weights and basis of the volume blocks, [q][Np]
  __shared__ double s_tabf[2][NFF * 4];   // ... and of the faces
  __shared__ double s_geo[EB * GEO_STRIDE];
  __shared__ double s_tau[EB * 4];
  __shared__ int s_conn[EB * 4];
  __shared__ int s_fmx[16 * Nfp];

  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int w = __builtin_amdgcn_readfirstlane(tid >> 6);  // wave-uniform: role branches are scalar
  const bool cw = w < RT;                                   // contraction wave (tile w) or helper
  const int K = a.K;
  const size_t KNp = (size_t)K * Np;

  if (MODE == 0 && !a.scal->active) {
    // Past tmax: the stage is the identity on u (k untouched).
    for (size_t i = (size_t)blockIdx.x * NT + tid; i < 5 * KNp; i += (size_t)gridDim.x * NT) a.u_out[i] = a.u_in[i];
    return;
  }
  const double dt = MODE == 0 ? a.scal->dt : 0.0;
  const double gamma = a.gamma;

  for (int i = tid; i < 16 * Nfp; i += NT) s_fmx[i] = a.fmx[i];
  {
    const double *pv = (const double *)(tab + TB::pv_off), *vv = (const double *)(tab + TB::vv_off);
    const double *pf = (const double *)(tab + TB::pf_off), *vf = (const double *)(tab + TB::vf_off);
    for (int i = tid; i < Np * 4; i += NT) {
      s_tabv[0][(i & 3) * Np + (i >> 2)] = pv[i];
      s_tabv[1][(i & 3) * Np + (i >> 2)] = vv[i];
    }
    for (int i = tid; i < NFF * 4; i += NT) {
      s_tabf[0][i] = pf[i];
      s_tabf[1][i] = vf[i];
    }
  }
  for (int i = 3 * Np * S::RS + tid; i < KV4 * S::RS; i += NT) s_vr[0][i] = s_vr[1][i] = 0.0f;  // padded rows stay zero
  const double *pv = s_tabv[0], *vv = s_tabv[1], *pf = s_tabf[0], *vf = s_tabf[1];

  const int grp = blockIdx.x % ngrp, gsl = blockIdx.x / ngrp, gsz = gridDim.x / ngrp;
  const int lo = (int)((long)nbatch * grp / ngrp), hi = (int)((long)nbatch * (grp + 1) / ngrp);

  // Staging values of a batch (node values, geometry, connectivity, tau),
  // loaded by the helper waves only: the contraction waves' registers hold
  // their operator rows (register loads there had to spill, and direct-to-LDS
  // copies issued by two waves were slower).  The helpers load the next
  // batch's values during the current batch's variable loop (the
  // connectivity after its first volume pass, the rest after its last one,
  // so that no volume pass holds them), and only their LDS writes remain at
  // the next stage-in.  Columns past the last element hold a state at rest.
  constexpr int NHW = S::NW - RT;                                      // helper waves
  constexpr int NHX = (EB * Np + 64 * NHW - 1) / (64 * NHW);           // node values per lane and variable
  constexpr int NHG = (EB * GEO_STRIDE + 64 * NHW - 1) / (64 * NHW);  // geometry values per lane
  const int hl = tid - 64 * RT, HT = 64 * NHW;                         // helper thread index and count
  double ld[5][NHX], lg[NHG];
  int lc = 0;       // helper wave 0: connectivity word of face `lane` of the batch
  double lt = 0.0;  // ... and its tau
  auto fetch = [&](int e0, int ne) {
#pragma unroll
    for (int v = 0; v < 5; v++) {
      const double *bv = a.u_in + v * KNp + (size_t)e0 * Np;
#pragma unroll
      for (int x = 0; x < NHX; x++) {
        const int i = hl + HT * x;
        ld[v][x] = i < ne * Np ? bv[i] : (v == 0 ? 1.0 : (v == 4 ? 2.5 : 0.0));
      }
    }
#pragma unroll
    for (int x = 0; x < NHG; x++) {
      const int i = hl + HT * x;
      lg[x] = i < ne * GEO_STRIDE ? a.geo[(size_t)e0 * GEO_STRIDE + i] : (i % GEO_STRIDE < 9 ? 0.0 : 1.0);
    }
  };
  auto fetch_conn = [&](int e0, int ne) {
    if (w == RT) lc = lane < ne * 4 ? a.conn[(size_t)e0 * 4 + lane] : 0;
  };
  // tau of the batch's faces (reads the connectivity word: called well after
  // fetch_conn, and before fetch)
  auto fetch_tau = [&](int e0, int ne) {
    if (w == RT)
      lt = lane < ne * 4 ? 0.25 * (a.maxv[(size_t)e0 * 4 + lane] + a.maxv[conn_nbr(lc) * 4 + conn_face(lc)]) *
                               a.tau_scale
                         : 0.0;
  };

  // Stages 1-3 of a batch, run by every wave (each role calls it at the same
  // point of its own loop; both roles pass the same barriers).  `helper` is
  // a constant at each call, so each role's copy holds only its own registers.
  auto stage_in = [&](int e0, int ne, const bool helper) {
    mx_barrier();  // the previous batch is done with every LDS buffer
    // ---- 1. the helpers' fetched values into LDS ----
    if (helper) {
#pragma unroll
      for (int x = 0; x < NHX; x++) {
        const int i = hl + HT * x, e = i / Np, j = i - e * Np;
        if (i < EB * Np)
#pragma unroll
          for (int v = 0; v < 5; v++) s_u[v][e * UST + j] = ld[v][x];
      }
#pragma unroll
      for (int x = 0; x < NHG; x++) {
        const int i = hl + HT * x;
        if (i < EB * GEO_STRIDE) s_geo[i] = lg[x];
      }
      if (w == RT) {
        s_conn[lane] = lc;
        s_tau[lane] = lt;
      }
    }
    mx_barrier();
    // ---- 2. surface fluxes of all five variables at every face-node pair,
    //      their least-squares face fits and the FP32 remainders.  Thread
    //      (face fi = (f, e), group s) takes face nodes jf = s + 8x; the
    //      fit sums meet across the eight groups of a face in registers ----
    {
      const int fi = tid >> 3, s8 = tid & 7, e = fi & 15, f = fi >> 4;
      const int cc = s_conn[e * 4 + f];
      const double *g = &s_geo[e * GEO_STRIDE + 9 + f * 4];
      const double tau = s_tau[e * 4 + f];
      double fl[JF][5];
      {
        double uP[JF][5];  // exterior traces, gathered together
#pragma unroll
        for (int x = 0; x < JF; x++) {
          const int jf = s8 + 8 * x;
          size_t pidx = 0;  // dead pairs read node 0 (unused)
          if (jf < Nfp && e < ne)
            pidx = (size_t)conn_nbr(cc) * Np + s_fmx[(conn_face(cc) * 4 + conn_code(cc)) * Nfp + jf];
#pragma unroll
          for (int v = 0; v < 5; v++) uP[x][v] = a.u_in[v * KNp + pidx];
        }
#pragma unroll
        for (int x = 0; x < JF; x++) {
          const int jf = s8 + 8 * x;
#pragma unroll
          for (int v = 0; v < 5; v++) fl[x][v] = 0.0;
          if (jf < Nfp && e < ne) {
            const int q = e * UST + s_fmx[(f * 4 + 3) * Nfp + jf];
            double uM[5];
#pragma unroll
            for (int v = 0; v < 5; v++) uM[v] = s_u[v][q];
            mx_surface_flux5(uM, uP[x], g, tau, a.flux, gamma, fl[x]);
          }
        }
      }
      double c[5][3] = {};
#pragma unroll
      for (int x = 0; x < JF; x++) {
        const int jf = s8 + 8 * x;
        if (jf < Nfp) {
          const double *P = &pf[(f * Nfp + jf) * 4];
#pragma unroll
          for (int v = 0; v < 5; v++)
#pragma unroll
            for (int qq = 0; qq < 3; qq++) c[v][qq] += P[qq] * fl[x][v];
        }
      }
#pragma unroll
      for (int v = 0; v < 5; v++)
#pragma unroll
        for (int qq = 0; qq < 3; qq++) c[v][qq] = sum8(c[v][qq]);
      if (s8 == 0)
#pragma unroll
        for (int v = 0; v < 5; v++)
#pragma unroll
          for (int qq = 0; qq < 3; qq++) s_c[v][(12 + f * 3 + qq) * EB + e] = c[v][qq];
#pragma unroll
      for (int x = 0; x < JF; x++) {
        const int jf = s8 + 8 * x;
        if (jf < Nfp) {
          const int m = f * Nfp + jf;
          const double *V = &vf[m * 4];
#pragma unroll
          for (int v = 0; v < 5; v++)
            s_sr[v][m * EB + e] =
                e < ne ? (float)(fl[x][v] - (V[0] * c[v][0] + V[1] * c[v][1] + V[2] * c[v][2])) : 0.0f;
        }
      }
    }
    // ---- 3. variable 0's volume pass, all eight slots on waves 0..3: the
    //      waves of faces 2 and 3 (4..7), whose exterior traces lie farther
    //      away in memory, finish the surface pass later ----
    if (w < 4) {
      mx_volume_pass<N>(0, w, lane, ne, &s_u[0][0], s_geo, pv, vv, gamma, s_c[0], s_vr[0]);
      mx_volume_pass<N>(0, w + 4, lane, ne, &s_u[0][0], s_geo, pv, vv, gamma, s_c[0], s_vr[0]);
    }
    mx_barrier();
  };

  if (!cw) {
    // ---- helper waves: their share of the stages and of every volume pass,
    //      and the next batch's staging loads ----
    if (lo + gsl < hi) {
      const int b = order ? order[lo + gsl] : lo + gsl;
      const int e0 = b * EB, ne = min(EB, K - e0);
      fetch_conn(e0, ne);
      fetch_tau(e0, ne);
      fetch(e0, ne);
    }
    for (int bq = lo + gsl; bq < hi; bq += gsz) {
      const int b = order ? order[bq] : bq;
      const int e0 = b * EB, ne = min(EB, K - e0);
      stage_in(e0, ne, true);
      // the next batch (after the last one, this one again: the fetches are
      // unconditional, so that no path carries their registers through the
      // volume passes)
      const int bn = bq + gsz < hi ? (order ? order[bq + gsz] : bq + gsz) : b;
      const int e1 = bn * EB, ne1 = min(EB, K - e1);
#pragma unroll 1
      for (int v = 0; v < 4; v++) {
        mx_volume_pass<N>(v + 1, w, lane, ne, &s_u[0][0], s_geo, pv, vv, gamma, s_c[v + 1], s_vr[(v + 1) & 1]);
        if constexpr (RT == 6)  // the slot of contraction wave w - 2 (see below)
          mx_volume_pass<N>(v + 1, w - 2, lane, ne, &s_u[0][0], s_geo, pv, vv, gamma, s_c[v + 1], s_vr[(v + 1) & 1]);
        if (v == 0) fetch_conn(e1, ne1);
        mx_barrier();
      }
      // (tau first: its wait for the connectivity word would also wait for
      // the node-value loads issued before it)
      fetch_tau(e1, ne1);
      fetch(e1, ne1);
      mx_barrier();  // the contraction waves' last variable
    }
    return;
  }

  // ---- contraction waves: this wave's operator rows (all k-steps) and (A V)
  //      rows, resident for the whole kernel ----
  float afr[KS];
  double avr[NC / 4];
#pragma unroll
  for (int x = 0; x < KS; x++) afr[x] = afrag[((size_t)w * KS + x) * 64 + lane];
#pragma unroll
  for (int x = 0; x < NC / 4; x++) avr[x] = avfrag[((size_t)w * (NC / 4) + x) * 64 + lane];
  const int kq = lane >> 4, ec = lane & 15;  // MFMA fragment row group and column

  for (int bq = lo + gsl; bq < hi; bq += gsz) {
    const int b = order ? order[bq] : bq;
    const int e0 = b * EB, ne = min(EB, K - e0);
    stage_in(e0, ne, false);
    // ---- 4. per variable: the contraction, the fit term, the update, then
    //      this wave's share of the next variable's volume pass ----
#pragma unroll 1
    for (int v = 0; v < 5; v++) {
      // the LSRK register of rows i0 + r, loaded ahead of the chain; a lane's
      // four rows are consecutive in memory, so whole rows go as two 16-byte
      // accesses (8-byte accesses scattered over 16 elements x 4 row groups
      // cost a memory transaction per lane, and every later wait of this
      // wave on a load also waits for its stores to drain)
      const int i0 = 16 * w + 4 * kq;
      const size_t g0 = v * KNp + (size_t)(e0 + ec) * Np + i0;
      const bool full = i0 + 3 < Np;
      double kold[4] = {0.0, 0.0, 0.0, 0.0};
      if (MODE == 0 && ec < ne) {
        if (full) {
          const f64x2u k01 = *reinterpret_cast<const f64x2u *>(a.k + g0);
          const f64x2u k23 = *reinterpret_cast<const f64x2u *>(a.k + g0 + 2);
          kold[0] = k01[0];
          kold[1] = k01[1];
          kold[2] = k23[0];
          kold[3] = k23[1];
        } else {
#pragma unroll
          for (int r = 0; r < 4; r++)
            if (i0 + r < Np) kold[r] = a.k[g0 + r];
        }
      }
      const float *vr = s_vr[v & 1];
      f32x4 acc = {0.f, 0.f, 0.f, 0.f};
      // (a scheduling fence every 8 k-steps bounds the LDS reads hoisted
      // ahead of the chain, and with them the live registers)
#pragma unroll
      for (int ks = 0; ks < KS; ks++) {
        const float bf = ks < KSV ? vr[(4 * ks + kq) * S::RS + ec] : s_sr[v][(4 * (ks - KSV) + kq) * EB + ec];
        acc = __builtin_amdgcn_mfma_f32_16x16x4f32(afr[ks], bf, acc, 0, 0, 0);
        if (ks % 8 == 7) __builtin_amdgcn_sched_barrier(0);
      }
      f64x4 acc64 = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
      for (int ks = 0; ks < NC / 4; ks++)
        acc64 = __builtin_amdgcn_mfma_f64_16x16x4f64(avr[ks], s_c[v][(4 * ks + kq) * EB + ec], acc64, 0, 0, 0);
      if (ec < ne) {
        double o1[4], o2[4];  // MODE 0: k and u_out; MODE 1: rhs
#pragma unroll
        for (int r = 0; r < 4; r++) {
          const double rhs = (double)acc[r] + acc64[r];
          if (MODE == 1) {
            o1[r] = rhs;
            o2[r] = 0.0;
          } else {
            o1[r] = a.rka * kold[r] + dt * rhs;
            o2[r] = s_u[v][ec * UST + min(i0 + r, Np - 1)] + a.rkb * o1[r];
          }
        }
        double *d1 = MODE == 1 ? a.rhs : a.k;
        if (full) {
          *reinterpret_cast<f64x2u *>(d1 + g0) = f64x2u{o1[0], o1[1]};
          *reinterpret_cast<f64x2u *>(d1 + g0 + 2) = f64x2u{o1[2], o1[3]};
          if (MODE == 0) {
            *reinterpret_cast<f64x2u *>(a.u_out + g0) = f64x2u{o2[0], o2[1]};
            *reinterpret_cast<f64x2u *>(a.u_out + g0 + 2) = f64x2u{o2[2], o2[3]};
          }
        } else {
#pragma unroll
          for (int r = 0; r < 4; r++)
            if (i0 + r < Np) {
              d1[g0 + r] = o1[r];
              if (MODE == 0) a.u_out[g0 + r] = o2[r];
            }
        }
      }
      // (at RT = 6, N = 6, contraction waves 4 and 5 share SIMDs 0 and 1 with
      // waves 0 and 1, whose chains take the MFMA pipe first: the helpers, on
      // SIMDs 2 and 3, run their volume-pass slots in the variable loop)
      if (v < 4 && (RT != 6 || w < 4))
        mx_volume_pass<N>(v + 1, w, lane, ne, &s_u[0][0], s_geo, pv, vv, gamma, s_c[v + 1], s_vr[(v + 1) & 1]);
      mx_barrier();  // remainder tile of v + 1 complete; the tile of v is free
    }
  }
}

}  // namespace dglm
