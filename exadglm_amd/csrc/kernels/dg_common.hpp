// Shared device-side definitions for the DGLM kernels (gfx950).
#pragma once
#include <hip/hip_runtime.h>

namespace dglm {

constexpr int NFACE = 4;

template <int N>
struct Ord {
  static constexpr int Np = (N + 1) * (N + 2) * (N + 3) / 6;
  static constexpr int Nfp = (N + 1) * (N + 2) / 2;
  static constexpr int NFF = NFACE * Nfp;
  static constexpr int KD = 3 * Np + NFF;  // stacked contraction length [WDr|WDs|WDt|LIFT]
};

// Per-step scalars that live in device memory so the step loop never reads
// back to the host (the reference copies dt to the host every step,
// euler_nopad_dglm.cu:79-96).
struct DevScalars {
  double t;        // start time of the current step
  double t_next;   // time after the current step (= committed time)
  double dt;       // step size of the current step
  double maxfs;    // max over faces of lambda * Fscale
  int tstep;       // completed steps after the current one
  int active;      // 0 once t_next >= tmax: stage kernels become no-ops
  int bad;         // non-finite dt seen
  int fstar;       // face attaining max(lambda * Fscale) (smallest index on ties)
  int clipped;     // 1 if dt of the current step was clipped to end at tmax
  int pad;
};

// Per-element geometry record: 9 metrics followed by 4x(nx,ny,nz,Fscale).
constexpr int GEO_STRIDE = 25;

// Packed face connectivity: (nbr << 4) | (nbr_face << 2) | code, where code
// 0..2 selects FmaskP[nbr_face][rot] and code 3 selects Fmask[nbr_face]
// (a boundary face whose exterior trace is the interior one).
__host__ __device__ inline int conn_nbr(int c) { return c >> 4; }
__host__ __device__ inline int conn_face(int c) { return (c >> 2) & 3; }
__host__ __device__ inline int conn_code(int c) { return c & 3; }

struct PhysParams {
  double gamma;
  double tau_scale;
  double cfl_fac;  // CFL / (2N+1)
  double tmax;
};

}  // namespace dglm
