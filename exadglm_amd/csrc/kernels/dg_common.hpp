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
// Reciprocal and square root for dependency chains that bound a kernel (the
// fused mixed-precision kernel's surface and volume passes; in the adjoint's
// pass A and the FP64 warp-specialised kernel they were measured no faster): the
// hardware estimates v_rcp_f64 and v_rsq_f64 refined by Newton steps (two
// r += r (1 - x r); for the root a coupled step on g = x y, h = y / 2 and a
// correction g += h (x - g^2)), within an ulp of the IEEE results and equal
// to them on every one of 4M sampled densities in [1e-3, 1e3] and squared
// sound speeds in [1e-6, 1e6]; the IEEE division and square-root sequences
// are chains about twice as long.  For positive normal arguments (densities,
// squared sound speeds); zero and NaN pass through the root.
__device__ inline double rcp_refined(double x) {
  double r = __builtin_amdgcn_rcp(x);
  double e = __builtin_fma(-x, r, 1.0);
  r = __builtin_fma(r, e, r);
  e = __builtin_fma(-x, r, 1.0);
  return __builtin_fma(r, e, r);
}
__device__ inline double sqrt_refined(double x) {
  const double y = __builtin_amdgcn_rsq(x);
  double g = x * y, h = 0.5 * y;
  const double r = __builtin_fma(-h, g, 0.5);
  g = __builtin_fma(g, r, g);
  h = __builtin_fma(h, r, h);
  g = __builtin_fma(__builtin_fma(-g, g, x), h, g);
  return x > 0.0 ? g : x;
}

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
