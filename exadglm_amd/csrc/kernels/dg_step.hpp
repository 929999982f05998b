// Per-step reductions that are plain (non-template) kernels: the device-side
// dt and the conserved integrals.  Included by solver.hip only, so each is
// defined in one translation unit.
//   calc_dt  dim3/src/cpp/euler_dglm.cpp:53-69, euler_time.cpp:49-57
#pragma once
#include "dg_common.hpp"

namespace dglm {

// dt = CFL/(2N+1) / max(lambda*Fscale), clipped so the run ends exactly at
// tmax (euler_time.cpp:49-57).  One block.
__global__ __launch_bounds__(1024) void finalize_dt_kernel(const double *__restrict__ partial,
                                                            const int *__restrict__ partial_idx, int n,
                                                            DevScalars *s, double cfl_fac, double tmax,
                                                            double *dt_hist, int *fstar_hist, int *clip_hist,
                                                            int hist_idx) {
  __shared__ double red[1024];
  __shared__ int red_i[1024];
  double m = 0.0;
  int mi = 0x7fffffff;
  for (int i = threadIdx.x; i < n; i += 1024) {
    const double o = partial[i];
    if (o > m || (o == m && partial_idx[i] < mi)) {
      m = o;
      mi = partial_idx[i];
    }
  }
  red[threadIdx.x] = m;
  red_i[threadIdx.x] = mi;
  __syncthreads();
  for (int w = 512; w > 0; w >>= 1) {
    if (threadIdx.x < w) {
      const double o = red[threadIdx.x + w];
      if (o > red[threadIdx.x] || (o == red[threadIdx.x] && red_i[threadIdx.x + w] < red_i[threadIdx.x])) {
        red[threadIdx.x] = o;
        red_i[threadIdx.x] = red_i[threadIdx.x + w];
      }
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const double maxfs = red[0];
    const double t = s->t_next;
    s->maxfs = maxfs;
    s->fstar = red_i[0];
    s->clipped = 0;
    if (t < tmax) {
      double dt = cfl_fac * (1.0 / maxfs);
      if (t + dt > tmax) {
        dt = tmax - t;
        s->clipped = 1;
      }
      s->t = t;
      s->dt = dt;
      s->t_next = t + dt;
      s->tstep += 1;
      s->active = 1;
      if (!(dt > 0.0) || !isfinite(dt)) s->bad = 1;
    } else {
      s->active = 0;
      s->dt = 0.0;
    }
    if (dt_hist) dt_hist[hist_idx] = s->active ? s->dt : 0.0;
    if (fstar_hist) fstar_hist[hist_idx] = s->fstar;
    if (clip_hist) clip_hist[hist_idx] = s->clipped;
  }
}

// Conserved integrals sum_e J_e w^T u_e (w = column sums of M), FP64: one
// thread per element, a fixed-order block reduction, partials per block.
__global__ __launch_bounds__(256) void integrals_kernel(int K, int Np, int nvar, const double *__restrict__ u,
                                                       const double *__restrict__ J,
                                                       const double *__restrict__ w, double *partial) {
  __shared__ double red[256];
  const int e = blockIdx.x * 256 + threadIdx.x;
  const size_t KNp = (size_t)K * Np;
  for (int v = 0; v < nvar; v++) {
    double s = 0.0;
    if (e < K) {
      const double *ue = u + v * KNp + (size_t)e * Np;
      for (int i = 0; i < Np; i++) s += w[i] * ue[i];
      s *= J[e];
    }
    red[threadIdx.x] = s;
    __syncthreads();
    for (int q = 128; q > 0; q >>= 1) {
      if (threadIdx.x < q) red[threadIdx.x] += red[threadIdx.x + q];
      __syncthreads();
    }
    if (threadIdx.x == 0) partial[(size_t)blockIdx.x * nvar + v] = red[0];
    __syncthreads();
  }
}

}  // namespace dglm
