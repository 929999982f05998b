// Warp-specialised FP64 stage kernel for the Euler equations (gfx950), and
// the LDS-staged wave-speed kernel.
//
// Stage kernel: one persistent 1024-thread workgroup per CU (16 waves, four
// per SIMD, so every wave fits in 128 VGPRs), batches of 16 elements in two
// LDS slots, two wave roles and two phases per batch (one barrier each):
//
//   phase 1  consumers contract the surface rows of batch b (the surface
//            fluxes, staged in one LDS buffer); producers copy batch b+1's
//            node values and geometry records straight into its LDS slot
//            (global_load_lds_dwordx4, no registers), wait for those and for
//            b+1's exterior traces (gathered a phase earlier), write tau;
//   phase 2  consumers contract the volume rows of batch b, apply the LSRK
//            update, then form b+1's pressure and contravariant velocity
//            V_d = sum_x g_dx v_x (the producers are the longer path of this
//            phase); producers form b+1's LM (or Rusanov) surface flux of all
//            five variables at every face node (written once per face node
//            into the surface buffer the consumers finished reading in
//            phase 1) and issue the gathers of b+2's exterior traces.
//
// Consumer wave (v, t) owns row tile t (16 nodes) of variable v and keeps its
// operator rows for the volume k-steps as v_mfma_f64_16x16x4f64 A-fragments
// in registers; those of the surface k-steps are read from LDS.  Its volume
// B-fragments are formed in registers: for variable v the metric flux is
// f_d = c V_d + p g_d,v-1  (c = rho, rho*u_v or E + p; the pressure term only
// for the momenta), one node's value feeding its three direction k-steps
// (contraction order k = d*Np4 + j).  The Np - 16*RTF rows that do not fill a
// tile are accumulated on the VALU, split between the t = 0 and t = 1 waves of
// the variable (operator rows in LDS, partial sums reduced across the four
// k-lanes with two shuffles), instead of padding a whole MFMA tile.  Batches are visited in an order table (Morton
// order of their centroids, built on the host) and the CUs of one XCD walk a
// contiguous stretch of it, so the traces shared by neighbouring batches are
// served from that XCD's L2.
//
// Reference semantics as in dg_fp64.hpp (euler_dglm.cpp:6-302,
// euler_time.cpp:15-38).
#pragma once
#include "dg_mfma.hpp"

namespace dglm {

template <int N>
struct WsShape {
  using O = Ord<N>;
  static constexpr int Np = O::Np, Nfp = O::Nfp, NFF = O::NFF;
  static constexpr int G = (Np + 3) / 4;  // node groups = k-steps per volume direction
  static constexpr int Np4 = 4 * G;
  static constexpr int KV = 3 * Np4;        // volume rows of the contraction
  static constexpr int KSS = (NFF + 3) / 4;  // surface k-steps
  static constexpr int NFF4 = 4 * KSS;
  static constexpr int KS = 3 * G + KSS;
  static constexpr int KDP = 4 * KS;
  static constexpr int RTF = Np / 16;         // full MFMA row tiles
  static constexpr int REM = Np - 16 * RTF;   // rows accumulated on the VALU
  static constexpr int NC = 5 * RTF;          // consumer waves
  static constexpr int NPW = 6;               // producer waves
  static constexpr int NW = NC + NPW;
  static constexpr int NT = 64 * NW;
  static constexpr int EB = 16;               // elements per batch (MFMA columns)
  static constexpr int EH = EB / 2;           // elements per DMA chunk (a half batch)
  static constexpr int NPAIR = EB * NFF4;     // face-node pairs of a batch
  static constexpr int PPW = (NPAIR + NPW - 1) / NPW;  // pairs per producer wave
  static constexpr int FS = EB + 1;           // row stride of the surface buffer: a producer wave stores
                                              // consecutive face nodes of one element, and the odd stride
                                              // spreads them over the LDS banks
};

// One batch's node data in LDS (two of these alternate).  Node arrays are
// element-major [element][node] (the global order, so a half batch is one
// contiguous DMA copy per variable); a B-fragment read of 16 elements then
// strides Np doubles, which spreads the lanes over distinct banks for odd Np.
template <int N>
struct WsSlot {
  using S = WsShape<N>;
  alignas(16) double u[5][S::EB * S::Np];
  alignas(16) double geo[S::EB * GEO_STRIDE];  // 9 metrics, 4 x (nx, ny, nz, Fscale)
  double p[S::EB * S::Np];
  double V[3][S::EB * S::Np];  // contravariant velocity
  double tau[S::EB * 4];
};

template <int N>
constexpr size_t ws_lds_bytes() {
  using S = WsShape<N>;
  return 2 * sizeof(WsSlot<N>) + 8 * 5 * S::NFF4 * S::FS + 8 * S::REM * S::KDP + 8 * S::RTF * S::KSS * 64 +
         4 * 16 * S::Nfp + 4 * 2 * S::EB * 4;
}

template <int N>
constexpr bool ws_available() {
  using S = WsShape<N>;
  return S::RTF >= 1 && S::NT <= 1024 && ws_lds_bytes<N>() <= 160 * 1024;
}

// Host-side: operator rows in the k order of the kernel.
//   afrag[(t*KS + ks)*64 + l] = A'[16t + l%16][4ks + l/16]   (t < RTF)
//   arem[r*KDP + k]            = A'[16*RTF + r][k]           (r < REM)
// with A'[i][d*Np4 + j] = op[(d*Np + j)*Np + i] (j < Np) and
// A'[i][KV + m] = op[(3Np + m)*Np + i] (m < NFF), zero elsewhere.
template <int N>
inline double ws_op(const double *opT, int i, int k) {
  using S = WsShape<N>;
  if (k < S::KV) {
    const int d = k / S::Np4, j = k - d * S::Np4;
    return j < S::Np ? opT[(size_t)(d * S::Np + j) * S::Np + i] : 0.0;
  }
  const int m = k - S::KV;
  return m < S::NFF ? opT[(size_t)(3 * S::Np + m) * S::Np + i] : 0.0;
}

template <int N>
inline void ws_arrange(const double *opT, double *afrag, double *arem) {
  using S = WsShape<N>;
  for (int t = 0; t < S::RTF; t++)
    for (int ks = 0; ks < S::KS; ks++)
      for (int l = 0; l < 64; l++)
        afrag[((size_t)t * S::KS + ks) * 64 + l] = ws_op<N>(opT, 16 * t + (l & 15), 4 * ks + (l >> 4));
  for (int r = 0; r < S::REM; r++)
    for (int k = 0; k < S::KDP; k++) arem[(size_t)r * S::KDP + k] = ws_op<N>(opT, 16 * S::RTF + r, k);
}

typedef __attribute__((address_space(3))) void lds_void;

// Copy `bytes` (a multiple of 8, at most 64*16*NI) from global src to LDS dst
// with one global_load_lds_dwordx4 per KiB: lane l of instruction q moves the
// 16 bytes at offset (64q + l)*16; lanes past the end are masked off (a final
// 8-byte tail is moved by a plain load and store).  dst 16-byte aligned and
// wave-uniform; a source that is not 16-byte aligned takes a plain copy.
template <int NI>
__device__ inline void dma_copy(const double *src, double *dst, int bytes, int lane) {
  if (((uintptr_t)src & 15) != 0) {
    for (int i = lane; i < (bytes >> 3); i += 64) dst[i] = src[i];
    return;
  }
  const int n16 = bytes >> 4;
#pragma unroll
  for (int q = 0; q < NI; q++) {
    const int i = 64 * q + lane;
    if (i < n16)
      __builtin_amdgcn_global_load_lds((const void *)(src + 2 * i), (lds_void *)(dst + 128 * q), 16, 0, 0);
  }
  if ((bytes & 15) && lane == 0) dst[2 * n16] = src[2 * n16];
}

// Role of hardware wave hw of the workgroup: r < NC is consumer (v, t) =
// (r / RTF, r % RTF), r >= NC producer r - NC.  Waves are dealt to the four
// SIMDs round-robin (SIMD hw % 4).  At N = 4 (16 waves: ten consumers, the
// t = 0 ones with two VALU remainder rows and the t = 1 ones with one, and six
// producers) the table puts three consumers and one producer on SIMDs 0 and 1
// and two of each on SIMDs 2 and 3, with the t = 0 consumers spread 1, 1, 2, 1,
// which evens out the FP64 work (MFMA and VALU share it) per SIMD; otherwise
// roles follow the wave index.
template <int N>
__device__ inline int ws_role(int hw) {
  using S = WsShape<N>;
  if constexpr (S::NW == 16 && S::RTF == 2 && S::REM > 0) {
    constexpr int tab[16] = {0, 2, 4, 8, 1, 5, 6, 9, 3, 7, 12, 14, 10, 11, 13, 15};
    int r = 0;
#pragma unroll
    for (int i = 0; i < 16; i++) r = i == hw ? tab[i] : r;
    return r;
  } else {
    return hw;
  }
}

__device__ inline int ws_batch(const int *order, int b) { return order ? order[b] : b; }

// First face-node pair of producer pw (pw = NPW: the end); producer pw takes
// pairs [ws_pair0(pw), ws_pair0(pw + 1)) of a batch.  A round of 64 pairs
// costs the same VALU issue whether its lanes are full or not, so at N = 4
// (960 pairs = 15 rounds) producers get whole rounds, ROUNDS_N4 per producer
// (the role table puts producers 0..5 on SIMDs 0, 1, 2, 2, 3, 3); otherwise
// an even split.
#ifndef WS_ROUNDS_N4
#define WS_ROUNDS_N4 2, 2, 3, 2, 3, 3
#endif
template <int N>
__host__ __device__ constexpr int ws_pair0(int pw) {
  using S = WsShape<N>;
  if constexpr (S::NPAIR == 15 * 64 && S::NPW == 6) {
    constexpr int r[6] = {WS_ROUNDS_N4};
    int c = 0;
    for (int i = 0; i < 6; i++) c += i < pw ? r[i] : 0;
    return 64 * c;
  } else {
    return pw * S::PPW < S::NPAIR ? pw * S::PPW : S::NPAIR;
  }
}

// Every producer's pairs fit its NRS register rounds, and together they cover
// the batch.
template <int N>
constexpr bool ws_pairs_fit() {
  using S = WsShape<N>;
  for (int pw = 0; pw < S::NPW; pw++)
    if (ws_pair0<N>(pw + 1) - ws_pair0<N>(pw) > 64 * ((S::PPW + 63) / 64)) return false;
  return ws_pair0<N>(0) == 0 && ws_pair0<N>(S::NPW) == S::NPAIR;
}

// Dirichlet flags of a producer lane's pairs: bit s set when pair s lies on
// a Dirichlet face (up[s] then holds the prescribed state).  Empty without
// Dirichlet faces, so that instantiation's register state is unchanged.
template <bool DIR>
struct WsDirBits {
  unsigned dm;
};
template <>
struct WsDirBits<false> {};

// Registers a producer wave carries from one staging step to the next.
template <int N, bool DIR>
struct WsProd : WsDirBits<DIR> {
  static constexpr int NRS = (WsShape<N>::PPW + 63) / 64;
  double up[NRS][5];        // exterior traces of this lane's face-node pairs
  double lam_own, lam_nbr;  // wave speeds of face `lane` of the batch (producer 0)
  int cn;                   // packed connectivity of face `lane` of the batch after (producer 0)
};

// Producer staging of batch b, spread over two batch periods:
//   ws_gather  (phase 2, a period ahead) register loads: the exterior trace
//              of face-node pairs P = PPW*pw + lane + 64 s of the batch
//              (element P / NFF4, face node P % NFF4, connectivity from
//              cb_cur), and for producer 0 the wave speeds of the batch's 64
//              faces and the connectivity of batch nb (< 0: none);
//   ws_dma     (phase 1) node values and geometry records straight into the
//              LDS slot (global_load_lds_dwordx4, no registers), two of the
//              twelve half-batch copies per producer, and the padding columns
//              of a short batch;
//   ws_land    (phase 1) wait for all of it; producer 0 writes tau and the
//              next connectivity.
// The other producers, and the consumers, read what one producer wrote only
// after the barrier that ends the phase.  DIR: the mesh has Dirichlet faces;
// their pairs load the prescribed state instead of the exterior trace.
template <int N, bool DIR>
__device__ inline void ws_gather(const StageArgs &a, const int *s_fmx, const int *cb_cur, int b, int nb, int pw,
                                 int lane_in, WsProd<N, DIR> &pr) {
  using S = WsShape<N>;
  constexpr int Np = S::Np, Nfp = S::Nfp, NFF = S::NFF, EB = S::EB, NFF4 = S::NFF4;
  constexpr int NRS = WsProd<N, DIR>::NRS;
  // lane-derived indices are the same every call: an opaque copy of the lane
  // id keeps them recomputed here rather than held in registers across the
  // batch loop
  int lane = lane_in;
  asm volatile("" : "+v"(lane));
  const int p0 = ws_pair0<N>(pw), np = ws_pair0<N>(pw + 1) - p0;  // this producer's pairs
  const int K = a.K;
  const size_t KNp = (size_t)K * Np;
  const int e0 = b * EB;
  const int ne = min(EB, K - e0);
  if (pw == 0) {
    pr.lam_own = 0.0;
    pr.lam_nbr = 0.0;
    if (lane < ne * 4) {
      const int cc = cb_cur[lane];
      pr.lam_own = a.maxv[(size_t)e0 * 4 + lane];
      pr.lam_nbr = a.maxv[(size_t)conn_nbr(cc) * 4 + conn_face(cc)];
    }
    pr.cn = 0;
    const int ne_n = nb >= 0 ? min(EB, K - nb * EB) : 0;
    if (lane < ne_n * 4) pr.cn = a.conn[(size_t)nb * EB * 4 + lane];
  }
  const double *ubase[5];
#pragma unroll
  for (int v = 0; v < 5; v++) ubase[v] = a.u_in + v * KNp;
  if constexpr (DIR) pr.dm = 0;
#pragma unroll
  for (int s = 0; s < NRS; s++) {
    const int pl = lane + 64 * s, P = p0 + pl;
    const int q = P / NFF4, m = P - q * NFF4;
    // dead pairs load node 0 (never used): no select on the loaded values,
    // so nothing here waits for the loads
    int pidx = 0;  // K*Np < 2^31 (checked on the host)
    int ds = -1, jd = 0;
    if (pl < np && q < ne && m < NFF) {
      const int f = m / Nfp, jf = m - f * Nfp;
      const int cc = cb_cur[q * 4 + f];
      pidx = conn_nbr(cc) * Np + s_fmx[(conn_face(cc) * 4 + conn_code(cc)) * Nfp + jf];
      if constexpr (DIR) {
        ds = a.dmap[(size_t)(e0 + q) * 4 + f];
        jd = jf;
      }
    }
    if constexpr (DIR) {
      if (ds >= 0) {
        pr.dm |= 1u << s;
#pragma unroll
        for (int v = 0; v < 5; v++) pr.up[s][v] = dirichlet_value(a, v, ds, jd, Nfp);
      } else {
#pragma unroll
        for (int v = 0; v < 5; v++) pr.up[s][v] = ubase[v][pidx];
      }
    } else {
#pragma unroll
      for (int v = 0; v < 5; v++) pr.up[s][v] = ubase[v][pidx];
    }
  }
}

template <int N>
__device__ inline void ws_dma(const StageArgs &a, WsSlot<N> &sl, int b, int pw, int lane) {
  using S = WsShape<N>;
  constexpr int Np = S::Np, EB = S::EB, EH = S::EH;
  const size_t KNp = (size_t)a.K * Np;
  constexpr int NIU = (EH * Np * 8 + 1023) / 1024;
  constexpr int NIG = (EH * GEO_STRIDE * 8 + 1023) / 1024;
#pragma unroll
  for (int x = 0; x < 2; x++) {
    const int c = 2 * pw + x;  // copy c: variable c / 2 (c < 10) or the geometry (c >= 10), half c % 2
    const int h = c & 1, el = EH * h, e0 = b * EB + el;
    const int ne = max(0, min(EH, a.K - e0));
    if (c < 10) {
      const int v = c >> 1;
      dma_copy<NIU>(a.u_in + v * KNp + (size_t)e0 * Np, &sl.u[v][el * Np], ne * Np * 8, lane);
      // columns past the last element: a harmless state (rho = 1, at rest)
      for (int i = ne * Np + lane; i < EH * Np; i += 64) sl.u[v][el * Np + i] = v == 0 ? 1.0 : 0.0;
    } else {
      dma_copy<NIG>(a.geo + (size_t)e0 * GEO_STRIDE, &sl.geo[el * GEO_STRIDE], ne * GEO_STRIDE * 8, lane);
      for (int i = ne * GEO_STRIDE + lane; i < EH * GEO_STRIDE; i += 64) sl.geo[el * GEO_STRIDE + i] = 0.0;
    }
  }
}

template <int N, bool DIR>
__device__ inline void ws_land(const StageArgs &a, WsSlot<N> &sl, int *cb_next, int pw, int lane,
                               const WsProd<N, DIR> &pr) {
  // ---- every load and DMA copy of this wave has landed ----
  __builtin_amdgcn_s_waitcnt(0);
  if (pw == 0) {
    sl.tau[lane] = 0.25 * (pr.lam_own + pr.lam_nbr) * a.tau_scale;
    cb_next[lane] = pr.cn;
  }
}

// Pressure and contravariant velocity V_d = sum_x g_dx v_x of every node of
// a staged batch; threads t = 0..nthr-1.
template <int N>
__device__ inline void ws_primitives(const StageArgs &a, WsSlot<N> &sl, int t, int nthr) {
  using S = WsShape<N>;
  constexpr int Np = S::Np;
  for (int o = t; o < S::EB * Np; o += nthr) {
    const int e = o / Np;
    const double r = sl.u[0][o], mx = sl.u[1][o], my = sl.u[2][o], mz = sl.u[3][o], E = sl.u[4][o];
    const double inv = 1.0 / r;
    sl.p[o] = (a.gamma - 1.0) * (E - 0.5 * (mx * mx + my * my + mz * mz) * inv);
    const double vx = mx * inv, vy = my * inv, vz = mz * inv;
    const double *g = &sl.geo[e * GEO_STRIDE];
#pragma unroll
    for (int d = 0; d < 3; d++) sl.V[d][o] = g[3 * d] * vx + g[3 * d + 1] * vy + g[3 * d + 2] * vz;
  }
}

// LM (or Rusanov) surface flux of the five variables times Fscale at this
// producer's face-node pairs of batch b, into fs ([variable][face node]
// [element], row stride FS); padding entries become zero.  On a Dirichlet
// pair the LM face state ub is the prescribed state (Rusanov: it is u+).
template <int N, bool DIR>
__device__ inline void ws_surface(const StageArgs &a, const WsSlot<N> &sl, double *fs, const int *s_fmx, int b,
                                  int pw, int lane_in, const WsProd<N, DIR> &pr) {
  using S = WsShape<N>;
  constexpr int Np = S::Np, Nfp = S::Nfp, NFF = S::NFF, EB = S::EB, NFF4 = S::NFF4;
  constexpr int NRS = WsProd<N, DIR>::NRS;
  int lane = lane_in;
  asm volatile("" : "+v"(lane));
  const int p0 = ws_pair0<N>(pw), np = ws_pair0<N>(pw + 1) - p0;  // this producer's pairs
  const int ne = min(EB, a.K - b * EB);
#pragma unroll
  for (int s = 0; s < NRS; s++) {
    const int pl = lane + 64 * s, P = p0 + pl;
    if (pl < np) {
      const int e = P / NFF4, m = P - e * NFF4;
      double fl[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
      if (e < ne && m < NFF) {
        const int f = m / Nfp, jf = m - f * Nfp;
        const int o = e * Np + s_fmx[(f * 4 + 3) * Nfp + jf];
        const double *g = &sl.geo[e * GEO_STRIDE + 9 + f * 4];
        double uM[5];
#pragma unroll
        for (int v = 0; v < 5; v++) uM[v] = sl.u[v][o];
        if (a.flux == 1) {
          euler_rusanov5(uM, pr.up[s], g, a.gamma, fl);
#pragma unroll
          for (int v = 0; v < 5; v++) fl[v] *= g[3];
        } else {
          const double tau = sl.tau[e * 4 + f];
          double ub[5];
#pragma unroll
          for (int v = 0; v < 5; v++) ub[v] = 0.5 * (uM[v] + pr.up[s][v]);
          if constexpr (DIR) {
            if ((pr.dm >> s) & 1u) {
#pragma unroll
              for (int v = 0; v < 5; v++) ub[v] = pr.up[s][v];
            }
          }
          const double inv = 1.0 / ub[0];
          const double p = (a.gamma - 1.0) * (ub[4] - 0.5 * (ub[1] * ub[1] + ub[2] * ub[2] + ub[3] * ub[3]) * inv);
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
      }
#pragma unroll
      for (int v = 0; v < 5; v++) fs[(v * NFF4 + m) * S::FS + e] = fl[v];
    }
    __builtin_amdgcn_sched_barrier(0);  // one face node at a time: bounds live registers
  }
}

// Remainder rows [lo, hi) of consumer tile t: the REM rows beyond the full
// tiles are split between the first two tiles of a variable (both read the
// same B-fragments), so that no single wave carries all of that VALU work
// into the surface phase.
template <int N>
__host__ __device__ constexpr int ws_rem_lo(int t) {
  using S = WsShape<N>;
  constexpr int h = S::RTF >= 2 ? (S::REM + 1) / 2 : S::REM;
  return t == 0 ? 0 : (t == 1 ? h : S::REM);
}
template <int N>
__host__ __device__ constexpr int ws_rem_hi(int t) {
  using S = WsShape<N>;
  return t == 0 ? ws_rem_lo<N>(1) : S::REM;
}

// Consumer wave (v, t): batches b, b+gsz, ... < hi from alternating slots, two
// barriers per batch.  It also accumulates the remainder rows [RLO, RHI) on
// the VALU (a separate instantiation per range, so every loop stays
// straight-line code), and forms the primitives of the next batch at the end
// of the volume phase.
template <int N, int MODE, int RLO, int RHI>
__device__ inline void ws_consume(const StageArgs &a, WsSlot<N> *sl, const double *fs, const double *s_arem,
                                  const double *s_afs, const double *__restrict__ afrag, const int *order, double dt,
                                  int b, int hi, int gsz, int v, int t, int lane) {
  using S = WsShape<N>;
  using P = PrecTraits<PREC_F64>;
  constexpr int Np = S::Np, EB = S::EB, G = S::G, Np4 = S::Np4, KV = S::KV, NFF4 = S::NFF4;
  constexpr int KS = S::KS, KSS = S::KSS, KDP = S::KDP, RTF = S::RTF;
  constexpr int REM = RHI - RLO;      // remainder rows of this wave
  constexpr int RA = REM > 0 ? REM : 1;
  constexpr int RR = (REM + 3) / 4;  // remainder rows per lane
  const int K = a.K;
  const size_t KNp = (size_t)K * Np;
  const int kk = lane >> 4, e = lane & 15;

  // operator rows of this wave for the volume k-steps, resident in registers
  // for the whole kernel (the empty asm makes them opaque, so they are never
  // re-loaded inside the loop); those of the surface k-steps are read from
  // LDS (s_afs) at each use
  double afr[3 * G];
#pragma unroll
  for (int ks = 0; ks < 3 * G; ks++) {
    afr[ks] = afrag[((size_t)t * KS + ks) * 64 + lane];
    asm volatile("" : "+v"(afr[ks]));
  }
  const double *afs = s_afs + (size_t)t * KSS * 64 + lane;
  const double *fsv = fs + v * NFF4 * S::FS;

  for (int it = 0; b < hi; b += gsz, it++) {
    const WsSlot<N> &s = sl[it & 1];
    const int e0 = ws_batch(order, b) * EB;
    const int ne = min(EB, K - e0);
    const bool live = e < ne;
    const size_t gbase = v * KNp + (size_t)(e0 + (live ? e : 0)) * Np;

    double kp[4], kpr[RR > 0 ? RR : 1];
    if (MODE == 0) {
#pragma unroll
      for (int r = 0; r < 4; r++) kp[r] = live ? a.k[gbase + 16 * t + 4 * r + kk] : 0.0;
#pragma unroll
      for (int x = 0; x < RR; x++) {
        const int rr = kk + 4 * x;
        kpr[x] = (live && rr < REM) ? a.k[gbase + 16 * RTF + RLO + rr] : 0.0;
      }
    }
    // The remainder-row operator values are the same every batch; an opaque
    // zero offset keeps them as LDS reads inside the loop instead of
    // REM*KS registers held across it.
    int zoff = 0;
    asm volatile("" : "+v"(zoff));
    const double *arw = s_arem + zoff + RLO * KDP + kk;

    typename P::Acc acc = {0, 0, 0, 0};
    double racc[RA];
#pragma unroll
    for (int r = 0; r < RA; r++) racc[r] = 0.0;

    // ---- phase 1: surface rows ----
#pragma unroll
    for (int ks = 0; ks < KSS; ks++) {
      const double bf = fsv[(4 * ks + kk) * S::FS + e];
      acc = P::mfma(afs[ks * 64], bf, acc);
#pragma unroll
      for (int r = 0; r < REM; r++) {
        racc[r] += arw[r * KDP + KV + 4 * ks] * bf;
        asm volatile("" : "+v"(racc[r]));
      }
      if (ks % 4 == 3) __builtin_amdgcn_sched_barrier(0);
    }
    __syncthreads();  // surface buffer free for the next batch's fluxes

    // ---- phase 2: volume rows, f_d = c V_d + p g_d,v-1 ----
    double gv[3] = {0.0, 0.0, 0.0};
    if (v >= 1 && v <= 3) {
#pragma unroll
      for (int d = 0; d < 3; d++) gv[d] = s.geo[e * GEO_STRIDE + 3 * d + v - 1];
    }
#pragma unroll
    for (int g = 0; g < G; g++) {
      const int j = 4 * g + kk;
      double f0 = 0.0, f1 = 0.0, f2 = 0.0;
      if (j < Np) {
        const int o = e * Np + j;
        const double V0 = s.V[0][o], V1 = s.V[1][o], V2 = s.V[2][o];
        if (v == 0) {
          const double c = s.u[0][o];
          f0 = c * V0;
          f1 = c * V1;
          f2 = c * V2;
        } else if (v == 4) {
          const double c = s.u[4][o] + s.p[o];
          f0 = c * V0;
          f1 = c * V1;
          f2 = c * V2;
        } else {
          const double c = s.u[v][o], p = s.p[o];
          f0 = c * V0 + p * gv[0];
          f1 = c * V1 + p * gv[1];
          f2 = c * V2 + p * gv[2];
        }
      }
      acc = P::mfma(afr[g], f0, acc);
      acc = P::mfma(afr[G + g], f1, acc);
      acc = P::mfma(afr[2 * G + g], f2, acc);
#pragma unroll
      for (int r = 0; r < REM; r++) {
        const double *ar = &arw[r * KDP + 4 * g];
        racc[r] += ar[0] * f0 + ar[Np4] * f1 + ar[2 * Np4] * f2;
        // materialise the partial sum here: left alone the compiler sinks
        // these FMAs to the epilogue and holds every operand until then
        asm volatile("" : "+v"(racc[r]));
      }
      // keep each node group's LDS reads next to its MFMAs (bounds the
      // registers the scheduler may spend on reads hoisted from later groups)
      __builtin_amdgcn_sched_barrier(0);
    }

    // ---- LSRK epilogue (element e, rows 16t + 4r + kk) ----
    if (live) {
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const int i = 16 * t + 4 * r + kk;
        const size_t gi = gbase + i;
        const double rhs = acc[r];
        if (MODE == 1) {
          a.rhs[gi] = rhs;
        } else {
          const double kn = a.rka * kp[r] + dt * rhs;
          a.k[gi] = kn;
          a.u_out[gi] = s.u[v][e * Np + i] + a.rkb * kn;
        }
      }
    }
    if (REM > 0) {
      // partial sums of the four k-lanes of each element
#pragma unroll
      for (int r = 0; r < REM; r++) {
        racc[r] += __shfl_xor(racc[r], 16, 64);
        racc[r] += __shfl_xor(racc[r], 32, 64);
      }
#pragma unroll
      for (int x = 0; x < RR; x++) {
        const int rr = kk + 4 * x;
        double rhs = 0.0;
#pragma unroll
        for (int r = 0; r < RA; r++)
          if (r == rr) rhs = racc[r];
        if (live && rr < REM) {
          const int i = 16 * RTF + RLO + rr;
          const size_t gi = gbase + i;
          if (MODE == 1) {
            a.rhs[gi] = rhs;
          } else {
            const double kn = a.rka * kpr[x] + dt * rhs;
            a.k[gi] = kn;
            a.u_out[gi] = s.u[v][e * Np + i] + a.rkb * kn;
          }
        }
      }
    }
    // primitives of the next batch (its node values landed in the surface
    // phase; the producers' surface fluxes do not read them)
    if (b + gsz < hi) ws_primitives<N>(a, sl[(it + 1) & 1], (v * RTF + t) * 64 + lane, S::NC * 64);
    __syncthreads();  // slot free for batch b + 2*gsz
  }
}

//   MODE 0: LSRK update; MODE 1: write the RHS only.
//   DIR: the mesh has Dirichlet faces (a.dmap / a.gbnd are read); the
//   instantiation without them carries none of that code.
// ngrp: number of XCD groups the grid is split into (8, or 1 for plain
// striding); gridDim.x must be a multiple of ngrp.
// order: batch visiting order (physical batch order[b] for schedule position b;
// null: b itself).
template <int N, int MODE, bool DIR>
__global__ __launch_bounds__(WsShape<N>::NT) void stage_ws_kernel(StageArgs a, const double *__restrict__ afrag,
                                                                   const double *__restrict__ arem,
                                                                   const int *__restrict__ order, int nbatch,
                                                                   int ngrp) {
  using S = WsShape<N>;
  constexpr int Np = S::Np, Nfp = S::Nfp, KDP = S::KDP, RTF = S::RTF, REM = S::REM, NC = S::NC, NT = S::NT;
  static_assert(ws_pairs_fit<N>(), "producer pair ranges exceed their register rounds");

  __shared__ WsSlot<N> sl[2];
  __shared__ double s_fs[5 * S::NFF4 * S::FS];  // surface fluxes of the batch being consumed
  __shared__ double s_arem[REM > 0 ? REM * KDP : 1];
  __shared__ double s_afs[RTF * S::KSS * 64];  // A-fragments of the surface k-steps, [tile][k-step][lane]
  __shared__ int s_fmx[16 * Nfp];
  __shared__ int s_cb[2][S::EB * 4];  // packed connectivity of the batch being staged / the next one

  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int w = __builtin_amdgcn_readfirstlane(ws_role<N>(tid >> 6));  // role, uniform: role branches are scalar
  const size_t KNp = (size_t)a.K * Np;

  if (MODE == 0 && !a.scal->active) {
    // Past tmax: the stage is the identity on u (k untouched).
    for (size_t i = (size_t)blockIdx.x * NT + tid; i < 5 * KNp; i += (size_t)gridDim.x * NT) a.u_out[i] = a.u_in[i];
    return;
  }
  const double dt = MODE == 0 ? a.scal->dt : 0.0;

  for (int i = tid; i < 16 * Nfp; i += NT) s_fmx[i] = a.fmx[i];
  for (int i = tid; i < REM * KDP; i += NT) s_arem[i] = arem[i];
  for (int i = tid; i < RTF * S::KSS * 64; i += NT) {
    const int tt = i / (S::KSS * 64), r = i - tt * S::KSS * 64;
    s_afs[i] = afrag[((size_t)tt * S::KS + 3 * S::G) * 64 + r];
  }

  // Batch schedule: group x = blockIdx % ngrp walks batches [lo, hi) with
  // the gridDim/ngrp workgroups of the group striding together.
  const int grp = blockIdx.x % ngrp, gsl = blockIdx.x / ngrp, gsz = gridDim.x / ngrp;
  const int lo = (int)((long)nbatch * grp / ngrp), hi = (int)((long)nbatch * (grp + 1) / ngrp);
  int b = lo + gsl;

  if (w == NC && b < hi) {
    // connectivity of the first batch (later batches' is prefetched)
    const int e0 = ws_batch(order, b) * S::EB;
    if (e0 * 4 + lane < a.K * 4) s_cb[0][lane] = a.conn[(size_t)e0 * 4 + lane];
  }
  __syncthreads();  // s_fmx, s_arem, s_afs, s_cb

  if (w >= NC) {
    // producer waves.  Staging call n (batch lo + gsl + n*gsz) reads the
    // connectivity buffer n & 1 and fills the other.  In the period of batch
    // b they copy batch b + gsz into the free slot (phase 1), then form its
    // primitives and surface fluxes and gather batch b + 2*gsz (phase 2).
    const int pw = w - NC;
    // physical batch of schedule position x (-1: past the range)
    auto pb = [&](int x) { return x < hi ? ws_batch(order, x) : -1; };
    WsProd<N, DIR> pr;
    if (b < hi) {
      ws_gather<N, DIR>(a, s_fmx, s_cb[0], pb(b), pb(b + gsz), pw, lane, pr);
      ws_dma<N>(a, sl[0], pb(b), pw, lane);
      ws_land<N, DIR>(a, sl[0], s_cb[1], pw, lane, pr);
    }
    __syncthreads();  // batch b staged
    if (b < hi) {
      ws_primitives<N>(a, sl[0], pw * 64 + lane, S::NPW * 64);
      ws_surface<N, DIR>(a, sl[0], s_fs, s_fmx, pb(b), pw, lane, pr);
    }
    if (b + gsz < hi) ws_gather<N, DIR>(a, s_fmx, s_cb[1], pb(b + gsz), pb(b + 2 * gsz), pw, lane, pr);
    __syncthreads();  // batch b's primitives and surface fluxes complete
    for (int it = 0; b < hi; b += gsz, it++) {
      const int nb = b + gsz, nnb = b + 2 * gsz;
      if (nb < hi) {
        ws_dma<N>(a, sl[(it + 1) & 1], pb(nb), pw, lane);
        ws_land<N, DIR>(a, sl[(it + 1) & 1], s_cb[it & 1], pw, lane, pr);
      }
      __syncthreads();  // consumers are done with the surface buffer; batch nb staged
      if (nb < hi) ws_surface<N, DIR>(a, sl[(it + 1) & 1], s_fs, s_fmx, pb(nb), pw, lane, pr);
      if (nnb < hi) ws_gather<N, DIR>(a, s_fmx, s_cb[it & 1], pb(nnb), pb(nnb + gsz), pw, lane, pr);
      __syncthreads();
    }
    return;
  }
  __syncthreads();  // batch b staged
  __syncthreads();  // batch b's primitives and surface fluxes complete
  const int v = w / RTF, t = w - v * RTF;
  constexpr int L0 = ws_rem_lo<N>(0), H0 = ws_rem_hi<N>(0), L1 = ws_rem_lo<N>(1), H1 = ws_rem_hi<N>(1);
  if (t == 0) {
    ws_consume<N, MODE, L0, H0>(a, sl, s_fs, s_arem, s_afs, afrag, order, dt, b, hi, gsz, v, t, lane);
  } else if constexpr (RTF <= 2) {
    ws_consume<N, MODE, L1, H1>(a, sl, s_fs, s_arem, s_afs, afrag, order, dt, b, hi, gsz, v, t, lane);
  } else {
    if (t == 1)
      ws_consume<N, MODE, L1, H1>(a, sl, s_fs, s_arem, s_afs, afrag, order, dt, b, hi, gsz, v, t, lane);
    else
      ws_consume<N, MODE, REM, REM>(a, sl, s_fs, s_arem, s_afs, afrag, order, dt, b, hi, gsz, v, t, lane);
  }
}

// --------------------------------------------------------------------------
// LDS-staged wave speed: a block converts the node values of WE elements to
// |v|^2 and |c|^2 with coalesced loads, then one thread per face takes the
// maxima over the face nodes (calc_maxv, euler_dglm.cpp:6-50,
// advec_dglm.cpp:6-35); the block keeps the largest lambda*Fscale with its
// face index (smallest index on ties).
// --------------------------------------------------------------------------
__host__ __device__ constexpr int wavespeed_elems(int Np) { return Np <= 35 ? 64 : (Np <= 84 ? 32 : 16); }

template <int N, int PHYS>
__global__ __launch_bounds__(256) void wavespeed_lds_kernel(int K, const double *__restrict__ u,
                                                            const double *__restrict__ vel,
                                                            const double *__restrict__ geo,
                                                            const int *__restrict__ fmx, double *__restrict__ maxv,
                                                            double *__restrict__ partial,
                                                            int *__restrict__ partial_idx, double gamma) {
  using O = Ord<N>;
  constexpr int Np = O::Np, Nfp = O::Nfp, WE = wavespeed_elems(O::Np);
  __shared__ double s_v2[WE * Np];
  __shared__ double s_c2[PHYS == PHYS_EULER ? WE * Np : 1];
  __shared__ int s_fm[4 * Nfp];
  __shared__ double red[256];
  __shared__ int red_i[256];
  const int tid = threadIdx.x;
  const size_t KNp = (size_t)K * Np;
  const int e0 = blockIdx.x * WE;
  const int ne = min(WE, K - e0);
  for (int i = tid; i < 4 * Nfp; i += 256) s_fm[i] = fmx[((i / Nfp) * 4 + 3) * Nfp + i % Nfp];
  for (int i = tid; i < ne * Np; i += 256) {
    const size_t idx = (size_t)e0 * Np + i;
    if (PHYS == PHYS_EULER) {
      const double u1 = u[idx], u2 = u[KNp + idx], u3 = u[2 * KNp + idx], u4 = u[3 * KNp + idx],
                   u5 = u[4 * KNp + idx];
      const double vx = u2 / u1, vy = u3 / u1, vz = u4 / u1;
      const double v2 = vx * vx + vy * vy + vz * vz;
      const double pres = (gamma - 1.0) * (u5 - 0.5 * u1 * v2);
      s_v2[i] = v2;
      s_c2[i] = fabs(gamma * pres / u1);
    } else {
      const double vx = vel[idx], vy = vel[KNp + idx], vz = vel[2 * KNp + idx];
      s_v2[i] = vx * vx + vy * vy + vz * vz;
    }
  }
  __syncthreads();
  const int fid = e0 * 4 + tid;
  double lfs = 0.0;
  if (tid < ne * 4) {
    const int e = tid >> 2, f = tid & 3;
    double mv2 = 0.0, mc2 = 0.0;
    for (int j = 0; j < Nfp; j++) {
      const int q = e * Np + s_fm[f * Nfp + j];
      const double v2 = s_v2[q];
      mv2 = v2 > mv2 ? v2 : mv2;
      if (PHYS == PHYS_EULER) {
        const double c2 = s_c2[q];
        mc2 = c2 > mc2 ? c2 : mc2;
      }
    }
    const double lam = PHYS == PHYS_EULER ? sqrt(mv2) + sqrt(mc2) : sqrt(mv2);
    maxv[fid] = lam;
    lfs = lam * geo[(size_t)(e0 + e) * GEO_STRIDE + 9 + f * 4 + 3];
    // A non-finite state (no maximum over a NaN) reaches dt as an infinite
    // wave speed: dt = 0 is flagged as bad on the device.
    if (lfs != lfs) lfs = INFINITY;
  }
  red[tid] = lfs;
  red_i[tid] = tid < ne * 4 ? fid : 0x7fffffff;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (tid < s) {
      const double o = red[tid + s];
      if (o > red[tid] || (o == red[tid] && red_i[tid + s] < red_i[tid])) {
        red[tid] = o;
        red_i[tid] = red_i[tid + s];
      }
    }
    __syncthreads();
  }
  if (tid == 0) {
    partial[blockIdx.x] = red[0];
    partial_idx[blockIdx.x] = red_i[0];
  }
}

}  // namespace dglm
