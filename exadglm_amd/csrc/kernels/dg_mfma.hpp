// MFMA stage kernel (gfx950): persistent workgroups, operator resident in
// registers, per-variable flux tiles in LDS, LSRK update as the epilogue.
//
// The stage contraction  rhs[v][e][i] = sum_k A[i][k] B_v[k][e]  with
// A = [WDr | WDs | WDt | -LIFT] (Np x KD) and B_v the stacked volume fluxes
// (rows d*Np + j) and LM surface fluxes (rows 3Np + m) of variable v is a
// tiny-M / huge-N product (update_rhs, dim3/src/cpp/euler_dglm.cpp:254-302;
// 4 cublasDgemmBatched calls in euler_nopad_dglm.cu:269-300).  Here:
//   * one workgroup = RT = ceil(Np/16) waves; wave w owns rows 16w..16w+15 and
//     keeps its KS = ceil(KD/4) A-fragments in registers for the whole kernel;
//   * a batch = 16 elements = the 16 MFMA columns; for each variable the
//     workgroup writes B_v into an LDS tile [KDP][16] (k-major, element
//     fastest: a B-fragment read is 64 consecutive words, conflict free);
//   * the D-fragment is consumed in registers by the LSRK update
//     k = a k + dt rhs, u_out = u_in + b k (FP64 in every precision).
//
// Precision modes:
//   PREC_F64  v_mfma_f64_16x16x4f64; D register r of lane l is row 4r + l/16,
//             column l%16.
//   PREC_F32  v_mfma_f32_16x16x4f32 on  B - c, where c is the least-squares
//             affine fit of B over the nodes of each (element, variable,
//             block) -- blocks are the three volume directions and the four
//             faces -- computed in FP64, and the exact FP64 term
//             sum_b sum_q c_bq (A_b V_b)_q  is added in the epilogue.  The
//             fluxes, pressure, wave speed and fit are FP64; only the
//             second-order remainder of each block is rounded to FP32.  D
//             register r of lane l is row 4(l/16) + r, column l%16.
//   PREC_BF   the same split, with the remainder and the operator each cut
//             into bf16 pieces x = x1 + x2 + x3 (x1 = bf16(x), x2 =
//             bf16(x - x1), ...) and contracted on v_mfma_f32_16x16x16_bf16:
//             a.bf_terms = 3 adds x1 y1 + x1 y2 + x2 y1 (the split-BF16
//             three-term product), 6 also x1 y3 + x2 y2 + x3 y1.  The leading
//             product and the cross terms go to separate FP32 accumulators,
//             summed in FP64 in the epilogue.  Lane l holds A[l%16][4(l/16)+q]
//             and B[4(l/16)+q][l%16], q = 0..3, of each 16-deep k-step; D as
//             PREC_F32.
#pragma once
#include <cmath>
#include <utility>

#include "dg_fp64.hpp"

namespace dglm {

typedef double f64x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef short s16x4 __attribute__((ext_vector_type(4)));

enum { PREC_F64 = 0, PREC_F32 = 1, PREC_BF = 2 };
constexpr int BF_PIECES = 3;  // bf16 pieces kept per operand value (6-term product)

// bf16 bits of x: round to nearest even of the FP32 value of x (finite x).
__host__ __device__ inline unsigned short bf16_bits(double x) {
  const float f = (float)x;
  unsigned u = __builtin_bit_cast(unsigned, f);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (unsigned short)(u >> 16);
}
__host__ __device__ inline double bf16_value(unsigned short h) {
  return (double)__builtin_bit_cast(float, (unsigned)h << 16);
}
// x = sum of BF_PIECES bf16 pieces plus a remainder below the last one's ulp.
__host__ __device__ inline void bf16_split(double x, unsigned short (&h)[BF_PIECES]) {
#pragma unroll
  for (int p = 0; p < BF_PIECES; p++) {
    h[p] = bf16_bits(x);
    x -= bf16_value(h[p]);
  }
}
constexpr int NBLK = 7;  // offset blocks: 3 volume directions + 4 faces

template <int N>
struct MfmaShape {
  using O = Ord<N>;
  static constexpr int RT = (O::Np + 15) / 16;  // row tiles = waves per workgroup
  static constexpr int KS = (O::KD + 3) / 4;    // k-steps of 4
  static constexpr int KDP = KS * 4;            // padded contraction length
  static constexpr int NT = RT * 64;            // threads per workgroup
  static constexpr int KS16 = (O::KD + 15) / 16;  // k-steps of 16 (split-BF16 contraction)
};

template <int PREC>
struct PrecTraits;
template <>
struct PrecTraits<PREC_F64> {
  typedef double T;
  typedef f64x4 Acc;
  __device__ static inline Acc mfma(double a, double b, Acc c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  __device__ static inline int row(int lane, int r) { return 4 * r + (lane >> 4); }
};
template <>
struct PrecTraits<PREC_F32> {
  typedef float T;
  typedef f32x4 Acc;
  __device__ static inline Acc mfma(float a, float b, Acc c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  __device__ static inline int row(int lane, int r) { return 4 * (lane >> 4) + r; }
};
template <>
struct PrecTraits<PREC_BF> {
  typedef float T;
  typedef f32x4 Acc;
  __device__ static inline Acc mfma(s16x4 a, s16x4 b, Acc c) {
    return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
  }
  __device__ static inline int row(int lane, int r) { return 4 * (lane >> 4) + r; }
};

// Host-side arrangement of the stacked operator into A-fragments:
// frag[(w*KS + ks)*64 + lane] = A[16w + lane%16][4ks + lane/16] (0 outside).
template <typename T>
inline void arrange_a_fragments(int Np, int KD, int RT, int KS, const double *opT /*[KD][Np]*/, T *frag) {
  for (int w = 0; w < RT; w++)
    for (int ks = 0; ks < KS; ks++)
      for (int l = 0; l < 64; l++) {
        int i = 16 * w + (l & 15), k = 4 * ks + (l >> 4);
        frag[((size_t)w * KS + ks) * 64 + l] = (T)((i < Np && k < KD) ? opT[(size_t)k * Np + i] : 0.0);
      }
}

// Host-side split-BF16 A-fragments: frag[((w*KS16 + ks)*BF_PIECES + p)*64 + l]
// holds piece p of A[16w + l%16][16ks + 4(l/16) + q], q = 0..3 (0 outside).
inline void arrange_a_bf16(int Np, int KD, int RT, int KS16, const double *opT /*[KD][Np]*/, s16x4 *frag) {
  for (int w = 0; w < RT; w++)
    for (int ks = 0; ks < KS16; ks++)
      for (int l = 0; l < 64; l++) {
        unsigned short h[4][BF_PIECES];
        for (int q = 0; q < 4; q++) {
          const int i = 16 * w + (l & 15), k = 16 * ks + 4 * (l >> 4) + q;
          bf16_split((i < Np && k < KD) ? opT[(size_t)k * Np + i] : 0.0, h[q]);
        }
        for (int p = 0; p < BF_PIECES; p++)
          for (int q = 0; q < 4; q++) frag[(((size_t)w * KS16 + ks) * BF_PIECES + p) * 64 + l][q] = (short)h[q][p];
      }
}

// Offset block of contraction row k (three volume directions, four faces).
__host__ __device__ inline int offset_block(int k, int Np, int Nfp) {
  return k < 3 * Np ? k / Np : 3 + (k - 3 * Np) / Nfp;
}

// Host-side tables of the affine offsets, from the node coordinates of one
// element (any affine image of the reference nodes spans the same space):
//   fitV[k*4 + q]   basis q at the node of row k: 1 and the coordinates
//                   relative to the block centroid (volume blocks), or 1 and
//                   two orthonormal in-plane coordinates (faces; q = 3 is 0);
//   fitP[k*4 + q]   least-squares weights, c_bq = sum_{k in b} fitP[k*4+q] B_k;
//   fitAV[(i*NBLK + b)*4 + q] = sum_{k in b} A[i][k] fitV[k*4+q]  (FP64).
// Returns false if a block's normal matrix is singular.
inline bool affine_fit_tables(int Np, int Nfp, const int *Fmask, const double *x, const double *y, const double *z,
                              const double *opT, double *fitV, double *fitP, double *fitAV) {
  const int KD = 3 * Np + 4 * Nfp;
  for (int b = 0; b < NBLK; b++) {
    const int k0 = b < 3 ? b * Np : 3 * Np + (b - 3) * Nfp, n = b < 3 ? Np : Nfp;
    auto node = [&](int j) { return b < 3 ? j : Fmask[(b - 3) * Nfp + j]; };
    double cx = 0, cy = 0, cz = 0;
    for (int j = 0; j < n; j++) {
      cx += x[node(j)];
      cy += y[node(j)];
      cz += z[node(j)];
    }
    cx /= n;
    cy /= n;
    cz /= n;
    double e1[3] = {1, 0, 0}, e2[3] = {0, 1, 0}, e3[3] = {0, 0, 1};
    int nb = 4;
    if (b >= 3) {
      // orthonormal in-plane axes from the face's node spread
      int ja = 0;
      double best = -1;
      for (int j = 0; j < n; j++) {
        const double dx = x[node(j)] - cx, dy = y[node(j)] - cy, dz = z[node(j)] - cz, d = dx * dx + dy * dy + dz * dz;
        if (d > best) best = d, ja = j;
      }
      double a[3] = {x[node(ja)] - cx, y[node(ja)] - cy, z[node(ja)] - cz};
      double la = std::sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]);
      for (double &v : a) v /= la;
      int jb = 0;
      best = -1;
      for (int j = 0; j < n; j++) {
        double d[3] = {x[node(j)] - cx, y[node(j)] - cy, z[node(j)] - cz};
        const double pr = d[0] * a[0] + d[1] * a[1] + d[2] * a[2];
        for (int c = 0; c < 3; c++) d[c] -= pr * a[c];
        const double q = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
        if (q > best) best = q, jb = j;
      }
      double bvec[3] = {x[node(jb)] - cx, y[node(jb)] - cy, z[node(jb)] - cz};
      const double pr = bvec[0] * a[0] + bvec[1] * a[1] + bvec[2] * a[2];
      for (int c = 0; c < 3; c++) bvec[c] -= pr * a[c];
      const double lb = std::sqrt(bvec[0] * bvec[0] + bvec[1] * bvec[1] + bvec[2] * bvec[2]);
      for (int c = 0; c < 3; c++) {
        e1[c] = a[c];
        e2[c] = bvec[c] / lb;
        e3[c] = 0.0;
      }
      nb = 3;
    }
    for (int j = 0; j < n; j++) {
      const double d[3] = {x[node(j)] - cx, y[node(j)] - cy, z[node(j)] - cz};
      double *V = &fitV[(size_t)(k0 + j) * 4];
      V[0] = 1.0;
      V[1] = d[0] * e1[0] + d[1] * e1[1] + d[2] * e1[2];
      V[2] = d[0] * e2[0] + d[1] * e2[1] + d[2] * e2[2];
      V[3] = nb == 4 ? d[0] * e3[0] + d[1] * e3[1] + d[2] * e3[2] : 0.0;
    }
    // normal matrix G = V^T V (nb x nb), inverse by Gauss-Jordan
    double G[4][8] = {};
    for (int p = 0; p < nb; p++) {
      for (int q = 0; q < nb; q++) {
        double sum = 0;
        for (int j = 0; j < n; j++) sum += fitV[(size_t)(k0 + j) * 4 + p] * fitV[(size_t)(k0 + j) * 4 + q];
        G[p][q] = sum;
      }
      G[p][4 + p] = 1.0;
    }
    for (int c = 0; c < nb; c++) {
      int piv = c;
      for (int r = c + 1; r < nb; r++)
        if (std::fabs(G[r][c]) > std::fabs(G[piv][c])) piv = r;
      if (std::fabs(G[piv][c]) < 1e-300) return false;
      for (int q = 0; q < 8; q++) std::swap(G[c][q], G[piv][q]);
      const double inv = 1.0 / G[c][c];
      for (int q = 0; q < 8; q++) G[c][q] *= inv;
      for (int r = 0; r < nb; r++)
        if (r != c) {
          const double f = G[r][c];
          for (int q = 0; q < 8; q++) G[r][q] -= f * G[c][q];
        }
    }
    for (int j = 0; j < n; j++)
      for (int q = 0; q < 4; q++) {
        double w = 0;
        if (q < nb)
          for (int p = 0; p < nb; p++) w += G[q][4 + p] * fitV[(size_t)(k0 + j) * 4 + p];
        fitP[(size_t)(k0 + j) * 4 + q] = w;
      }
  }
  for (int i = 0; i < Np; i++)
    for (int b = 0; b < NBLK; b++)
      for (int q = 0; q < 4; q++) {
        const int k0 = b < 3 ? b * Np : 3 * Np + (b - 3) * Nfp, n = b < 3 ? Np : Nfp;
        double sum = 0;
        for (int k = k0; k < k0 + n; k++) sum += opT[(size_t)k * Np + i] * fitV[(size_t)k * 4 + q];
        fitAV[((size_t)i * NBLK + b) * 4 + q] = sum;
      }
  (void)KD;
  return true;
}

// Volume flux rows of variable v at node q: (fR, fS, fT) = metric . (F, G, H).
// u[0..4] point at the element's node values of each conserved variable
// (only u[0] for advection), vel[0..2] at the advection velocity.
template <int PHYS>
__device__ inline void volume_flux_v(int v, const double *const u[5], const double *const vel[3], int q,
                                     const double *g, double gamma, double &f0, double &f1, double &f2) {
  double F, G, H;
  if (PHYS == PHYS_EULER) {
    const double r = u[0][q], mx = u[1][q], my = u[2][q], mz = u[3][q], E = u[4][q];
    const double inv = 1.0 / r;
    const double p = (gamma - 1.0) * (E - 0.5 * (mx * mx + my * my + mz * mz) * inv);
    const double vx = mx * inv, vy = my * inv, vz = mz * inv;
    if (v == 0) {
      F = mx;
      G = my;
      H = mz;
    } else {
      // momentum component (v = 1..3) or total enthalpy density (v = 4)
      // times velocity, plus the pressure on the diagonal
      const double c = v == 4 ? E + p : (v == 1 ? mx : v == 2 ? my : mz);
      F = c * vx + (v == 1 ? p : 0.0);
      G = c * vy + (v == 2 ? p : 0.0);
      H = c * vz + (v == 3 ? p : 0.0);
    }
  } else {
    const double uu = u[0][q];
    F = vel[0][q] * uu;
    G = vel[1][q] * uu;
    H = vel[2][q] * uu;
  }
  f0 = g[0] * F + g[1] * G + g[2] * H;
  f1 = g[3] * F + g[4] * G + g[5] * H;
  f2 = g[6] * F + g[7] * G + g[8] * H;
}

//   MODE 0: LSRK update; MODE 1: write the RHS only.
template <int N, int PHYS, int MODE, int PREC>
__global__ __launch_bounds__(MfmaShape<N>::NT) void stage_mfma_kernel(StageArgs a, const void *__restrict__ afrag_v,
                                                                      const double *__restrict__ fit,
                                                                      int nbatch) {
  using O = Ord<N>;
  using S = MfmaShape<N>;
  using P = PrecTraits<PREC>;
  typedef typename P::T T;
  constexpr int Np = O::Np, Nfp = O::Nfp, NFF = O::NFF, KD = O::KD;
  constexpr int KS = S::KS, KDP = S::KDP, NT = S::NT;
  constexpr int NV = NVar<PHYS>::value;
  constexpr int EB = 16;
  constexpr int NVEL = PHYS == PHYS_ADV ? 3 : 1;
  constexpr bool OFFS = PREC != PREC_F64;
  constexpr bool BF = PREC == PREC_BF;
  constexpr int KS16 = S::KS16;

  __shared__ double s_u[NV][EB * Np];
  __shared__ T s_f[BF ? 1 : KDP * EB];
  // split-BF16 remainder tile: piece p of rows 16ks + 4q + 0..3 of column e
  // at s_fb[p][(4ks + q)*EB + e]
  __shared__ s16x4 s_fb[BF ? BF_PIECES : 1][BF ? 4 * KS16 * EB : 1];
  __shared__ double s_vel[NVEL][PHYS == PHYS_ADV ? EB * Np : 1];
  __shared__ double s_geo[EB * GEO_STRIDE];
  __shared__ double s_tau[EB * 4];
  __shared__ double s_b64[OFFS ? KDP * EB : 1];    // FP64 flux tile before the fit is removed
  __shared__ double s_c[OFFS ? NBLK * 4 * EB : 1];  // affine-fit coefficients [block][q][element]
  // fit = [fitV (KD x 4) | fitP (KD x 4) | fitAV (Np x NBLK x 4)]
  const double *fitV = fit, *fitP = fit + 4 * KD, *fitAV = fit + 8 * KD;
  __shared__ int s_conn[EB * 4];
  __shared__ int s_dm[EB * 4];  // Dirichlet slot of each face (-1: none)
  __shared__ int s_fmx[16 * Nfp];

  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int w = tid >> 6;
  const int K = a.K;
  const size_t KNp = (size_t)K * Np;
  const T *afrag = (const T *)afrag_v;

  if (MODE == 0 && !a.scal->active) {
    // Past tmax: the stage is the identity on u (k untouched).
    for (size_t i = (size_t)blockIdx.x * NT + tid; i < NV * KNp; i += (size_t)gridDim.x * NT) a.u_out[i] = a.u_in[i];
    return;
  }
  const double dt = MODE == 0 ? a.scal->dt : 0.0;

  // Operator rows of this wave, resident for the whole kernel.
  T afr[BF ? 1 : KS];
  s16x4 afb[BF ? KS16 : 1][BF_PIECES];
  if constexpr (BF) {
    const s16x4 *ab = (const s16x4 *)afrag_v;
#pragma unroll
    for (int ks = 0; ks < KS16; ks++)
#pragma unroll
      for (int p = 0; p < BF_PIECES; p++) afb[ks][p] = ab[(((size_t)w * KS16 + ks) * BF_PIECES + p) * 64 + lane];
  } else {
#pragma unroll
    for (int ks = 0; ks < KS; ks++) afr[ks] = afrag[((size_t)w * KS + ks) * 64 + lane];
  }

  for (int i = tid; i < 16 * Nfp; i += NT) s_fmx[i] = a.fmx[i];
  if constexpr (BF) {
    // padded rows stay zero
    for (int i = tid; i < BF_PIECES * 4 * KS16 * EB; i += NT) s_fb[i / (4 * KS16 * EB)][i % (4 * KS16 * EB)] = s16x4{0, 0, 0, 0};
  } else {
    for (int i = KD * EB + tid; i < KDP * EB; i += NT) s_f[i] = (T)0;  // padded rows stay zero
  }

  // LM surface flux of variable v at face node m of batch element e.
  auto lm_flux = [&](int v, int e, int m) -> double {
    const int f = m / Nfp, j = m - f * Nfp;
    const int c = s_conn[e * 4 + f];
    const size_t pidx = (size_t)conn_nbr(c) * Np + s_fmx[(conn_face(c) * 4 + conn_code(c)) * Nfp + j];
    const int mloc = e * Np + s_fmx[(f * 4 + 3) * Nfp + j];
    const double *g = &s_geo[e * GEO_STRIDE + 9 + f * 4];
    const double tau = s_tau[e * 4 + f];
    const int ds = s_dm[e * 4 + f];
    if (PHYS == PHYS_EULER) {
      double ub[5];
      if (a.flux == 1) {
        double uM[5], uP[5];
#pragma unroll
        for (int q = 0; q < 5; q++) {
          uM[q] = s_u[q][mloc];
          uP[q] = ds >= 0 ? dirichlet_value(a, q, ds, j, Nfp) : a.u_in[q * KNp + pidx];
        }
        return euler_rusanov(v, uM, uP, g, a.gamma) * g[3];
      }
#pragma unroll
      for (int q = 0; q < 5; q++)
        ub[q] = ds >= 0 ? dirichlet_value(a, q, ds, j, Nfp) : 0.5 * (s_u[q][mloc] + a.u_in[q * KNp + pidx]);
      const double inv = 1.0 / ub[0];
      const double p = (a.gamma - 1.0) * (ub[4] - 0.5 * (ub[1] * ub[1] + ub[2] * ub[2] + ub[3] * ub[3]) * inv);
      const double vn = (g[0] * ub[1] + g[1] * ub[2] + g[2] * ub[3]) * inv;  // n . velocity
      double nF;
      if (v == 0) nF = g[0] * ub[1] + g[1] * ub[2] + g[2] * ub[3];
      else if (v == 4) nF = (ub[4] + p) * vn;
      else nF = ub[v] * vn + g[v - 1] * p;
      return (nF - tau * (ub[v] - s_u[v][mloc])) * g[3];
    } else {
      const double uM = s_u[0][mloc];
      const double uP = ds >= 0 ? dirichlet_value(a, 0, ds, j, Nfp) : a.u_in[pidx];
      const double ub = ds >= 0 && a.flux != 1 ? uP : 0.5 * (uM + uP);
      if (a.flux == 1) {
        const double vn = g[0] * s_vel[0][mloc] + g[1] * s_vel[1][mloc] + g[2] * s_vel[2][mloc];
        return (vn * ub + 0.5 * fabs(vn) * (uM - uP)) * g[3];
      }
      const double nF = g[0] * (s_vel[0][mloc] * ub) + g[1] * (s_vel[1][mloc] * ub) + g[2] * (s_vel[2][mloc] * ub);
      return (nF - tau * (ub - uM)) * g[3];
    }
  };

  for (int b = blockIdx.x; b < nbatch; b += gridDim.x) {
    const int e0 = b * EB;
    const int ne = min(EB, K - e0);
    __syncthreads();  // previous batch finished with s_u / s_geo
    for (int i = tid; i < ne * GEO_STRIDE; i += NT) s_geo[i] = a.geo[(size_t)e0 * GEO_STRIDE + i];
    for (int i = tid; i < ne * 4; i += NT) {
      const int c = a.conn[(size_t)e0 * 4 + i];
      s_conn[i] = c;
      s_tau[i] = 0.25 * (a.maxv[(size_t)e0 * 4 + i] + a.maxv[conn_nbr(c) * 4 + conn_face(c)]) * a.tau_scale;
      s_dm[i] = a.dmap ? a.dmap[(size_t)e0 * 4 + i] : -1;
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
          const double *uu[5], *vv[3];
          for (int q = 0; q < 5; q++) uu[q] = s_u[q < NV ? q : 0] + e * Np;
          for (int q = 0; q < 3; q++) vv[q] = s_vel[q < NVEL ? q : 0] + e * Np;
          volume_flux_v<PHYS>(v, uu, vv, j, &s_geo[e * GEO_STRIDE], a.gamma, f0, f1, f2);
        }
        if constexpr (OFFS) {
          s_b64[j * EB + e] = f0;
          s_b64[(Np + j) * EB + e] = f1;
          s_b64[(2 * Np + j) * EB + e] = f2;
        } else {
          s_f[j * EB + e] = f0;
          s_f[(Np + j) * EB + e] = f1;
          s_f[(2 * Np + j) * EB + e] = f2;
        }
      }
      // ---- B_v tile: LM surface rows (3Np + m) ----
      for (int idx = tid; idx < NFF * EB; idx += NT) {
        const int m = idx >> 4, e = idx & 15;
        double fl = 0.0;
        if (e < ne) fl = lm_flux(v, e, m);
        if constexpr (OFFS)
          s_b64[(3 * Np + m) * EB + e] = fl;
        else
          s_f[(3 * Np + m) * EB + e] = fl;
      }
      __syncthreads();
      if constexpr (OFFS) {
        // least-squares affine fit per (block, element): c_bq = sum_k fitP[k][q] B_k.
        // Wave w takes blocks w, w + RT, ...; lane (element e, quarter s) sums
        // rows k0 + s, k0 + s + 4, ... of the block for all four q, and two
        // shuffles combine the quarters (a fixed order: deterministic).
        for (int bb = w; bb < NBLK; bb += S::RT) {
          const int k0 = bb < 3 ? bb * Np : 3 * Np + (bb - 3) * Nfp, k1 = bb < 3 ? k0 + Np : k0 + Nfp;
          const int e = lane & 15, sq = lane >> 4;
          double c0 = 0.0, c1 = 0.0, c2 = 0.0, c3 = 0.0;
          for (int k = k0 + sq; k < k1; k += 4) {
            const double bv = s_b64[k * EB + e];
            const double *pk = &fitP[k * 4];
            c0 += pk[0] * bv;
            c1 += pk[1] * bv;
            c2 += pk[2] * bv;
            c3 += pk[3] * bv;
          }
#pragma unroll
          for (int o = 16; o <= 32; o <<= 1) {
            c0 += __shfl_xor(c0, o, 64);
            c1 += __shfl_xor(c1, o, 64);
            c2 += __shfl_xor(c2, o, 64);
            c3 += __shfl_xor(c3, o, 64);
          }
          if (sq == 0) {
            s_c[(bb * 4 + 0) * EB + e] = c0;
            s_c[(bb * 4 + 1) * EB + e] = c1;
            s_c[(bb * 4 + 2) * EB + e] = c2;
            s_c[(bb * 4 + 3) * EB + e] = c3;
          }
        }
        __syncthreads();
        // the remainder B - fit, rounded to the MFMA precision
        for (int idx = tid; idx < KD * EB; idx += NT) {
          const int k = idx >> 4, e = idx & 15;
          const double *cb = &s_c[offset_block(k, Np, Nfp) * 4 * EB + e];
          const double *V = &fitV[k * 4];
          const double fv = cb[0] * V[0] + cb[EB] * V[1] + cb[2 * EB] * V[2] + cb[3 * EB] * V[3];
          if constexpr (BF) {
            unsigned short h[BF_PIECES];
            bf16_split(s_b64[idx] - fv, h);
            const int pos = ((k >> 4) * 4 + ((k >> 2) & 3)) * EB + e;
#pragma unroll
            for (int p = 0; p < BF_PIECES; p++) s_fb[p][pos][k & 3] = (short)h[p];
          } else {
            s_f[idx] = (T)(s_b64[idx] - fv);
          }
        }
        __syncthreads();
      }

      // ---- contraction on MFMA ----
      typename P::Acc acc = {0, 0, 0, 0}, acl = {0, 0, 0, 0};
      if constexpr (BF) {
        // cross terms into acl (smallest first), the leading product into acc
        const bool six = a.bf_terms == 6;
#pragma unroll
        for (int ks = 0; ks < KS16; ks++) {
          const int pos = (ks * 4 + (lane >> 4)) * EB + (lane & 15);
          const s16x4 b0 = s_fb[0][pos], b1 = s_fb[1][pos];
          if (six) {
            const s16x4 b2 = s_fb[2][pos];
            acl = P::mfma(afb[ks][2], b0, acl);
            acl = P::mfma(afb[ks][1], b1, acl);
            acl = P::mfma(afb[ks][0], b2, acl);
          }
          acl = P::mfma(afb[ks][1], b0, acl);
          acl = P::mfma(afb[ks][0], b1, acl);
          acc = P::mfma(afb[ks][0], b0, acc);
        }
      } else {
#pragma unroll
        for (int ks = 0; ks < KS; ks++) {
          const T bf = s_f[(ks * 4 + (lane >> 4)) * EB + (lane & 15)];
          acc = P::mfma(afr[ks], bf, acc);
        }
      }

      // ---- LSRK epilogue (FP64) for element lane%16 ----
      const int e = lane & 15;
      if (e < ne) {
#pragma unroll
        for (int r = 0; r < 4; r++) {
          const int i = 16 * w + P::row(lane, r);
          if (i < Np) {
            double rhs = (double)acc[r];
            if (BF) rhs += (double)acl[r];
            if (OFFS) {
              // exact FP64 contraction of the fit: sum_b sum_q c_bq (A_b V_b)_q
              const double *av = &fitAV[(size_t)i * NBLK * 4];
#pragma unroll
              for (int bq = 0; bq < NBLK * 4; bq++) rhs += s_c[bq * EB + e] * av[bq];
            }
            const size_t gi = v * KNp + (size_t)(e0 + e) * Np + i;
            if (MODE == 1) {
              a.rhs[gi] = rhs;
            } else {
              const double kn = a.rka * a.k[gi] + dt * rhs;
              a.k[gi] = kn;
              a.u_out[gi] = s_u[v][e * Np + i] + a.rkb * kn;
            }
          }
        }
      }
      __syncthreads();  // s_f and s_c are rewritten for the next variable
    }
  }
}

}  // namespace dglm
