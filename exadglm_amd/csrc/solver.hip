// Solver context and C ABI of libexadglm_hip.so (see include/exadglm.h).
//
// One context = one single-partition problem resident in HBM:
//   state u[2] (ping-pong between stages), LSRK register k, per-element
//   geometry records, packed face connectivity, the stacked operator, per-face
//   wave speeds and the device-side step scalars (t, dt, tstep).
// Every launch goes to the context's stream; nothing synchronises inside a
// step (the reference blocks on a D2H copy of dt each step,
// euler_nopad_dglm.cu:79-96).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/exadglm.h"
#include "host/datain.h"
#include "host/mesh.h"
#include "host/refelem.h"
#include "kernels/dg_fp64.hpp"
#include "kernels/dg_step.hpp"
#include "kernels/dg_mfma.hpp"
#include "kernels/dg_ws.hpp"
#include "kernels/dg_mx.hpp"
#include "kernels/dg_adjoint.hpp"
#include "kernels/dg_adjoint_mfma.hpp"

namespace dglm {
// ws_dirichlet.hip: the warp-specialised stage kernel on meshes with Dirichlet
// faces (false: not built for this order).
bool launch_ws_dirichlet(int N, int mode, int grid, hipStream_t stream, const StageArgs &a, const double *afrag,
                         const double *arem, const int *order, int nbatch, int groups);
}  // namespace dglm

#define DGLM_NTIMERS 3  // wave speed, dt finalize, stage

namespace {
thread_local std::string g_err;

// Allocation faults requested by dglm_set_alloc_faults (a testing aid: the
// next n device allocations fail as if the device were out of memory).
std::atomic<int> g_alloc_faults{0};

int fail(int code, const std::string &msg) {
  g_err = msg;
  return code;
}

#define HIP_TRY(call)                                                                          \
  do {                                                                                         \
    hipError_t e_ = (call);                                                                    \
    if (e_ != hipSuccess)                                                                      \
      return fail(DGLM_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_));           \
  } while (0)

// Low-storage RK4(5) of Carpenter & Kennedy (1994), 2N-storage form
// (coefficients as in dim3/src/cpp/euler_data.cpp:41-60).
const double RKA[5] = {0.0, -567301805773.0 / 1357537059087.0, -2404267990393.0 / 2016746695238.0,
                       -3550918686646.0 / 2091501179385.0, -1275806237668.0 / 842570457699.0};
const double RKB[5] = {1432997174477.0 / 9575080441755.0, 5161836677717.0 / 13612068292357.0,
                       1720146321549.0 / 2090206949498.0, 3134564353537.0 / 4481467310338.0,
                       2277821191437.0 / 14882151754819.0};
}  // namespace

using namespace dglm;

struct dglm_ctx {
  dglm_params p{};
  int K = 0, Np = 0, Nfp = 0, NFF = 0, KD = 0, nvar = 0;
  int dev = 0;
  hipStream_t stream = nullptr;
  double *u[2] = {nullptr, nullptr};
  double *k = nullptr, *rhs = nullptr, *vel = nullptr, *geo = nullptr, *op = nullptr, *maxv = nullptr,
         *partial = nullptr, *dt_hist = nullptr;
  double *scratch = nullptr;  // this device's copy of the allocation-free scratch array
  int dt_hist_cap = 0;
  int *conn = nullptr, *fmx = nullptr, *partial_idx = nullptr;
  DevScalars *scal = nullptr;
  int cur = 0;
  int nblk_ws = 0;
  int variant = DGLM_STAGE_AUTO;  // requested stage kernel
  void *afrag = nullptr;          // operator A-fragments for the MFMA kernel (FP64 or FP32)
  double *rowsum = nullptr;       // FP64 affine-offset tables of the mixed-precision contraction
  int mfma_grid = 0;              // persistent grid size (0 = not yet sized)
  double *afrag_ws = nullptr;     // operator rows in the warp-specialised kernel's k order
  double *arem_ws = nullptr;      // rows left to the VALU by that kernel
  int *ws_order = nullptr;        // its batch visiting order (null: mesh order)
  unsigned char *mx_tab = nullptr;  // tables of the mixed-precision Euler kernel (null: not built)
  int mx_grid = 0, mx_groups = 1;    // its persistent grid and XCD groups
  int ws_grid = 0, ws_groups = 1;  // its persistent grid and XCD groups
  struct dglm_adj_tape *tape = nullptr;  // adjoint buffers and recorded trajectory
  bool use_graph = true;                 // replay two-step hipGraphs in dglm_step
  hipGraphExec_t graph[2] = {nullptr, nullptr};  // keyed by the state buffer the step starts from
  double *d_J = nullptr, *d_w = nullptr, *d_ipart = nullptr;  // conservation sums
  bool adj_exact = true;          // include the tau(u) and dt(u) sensitivities (Euler)
  int adj_keep_max = -1;          // segments kept whole on the tape (-1: as many as half the free HBM holds)
  std::vector<int> h_conn;  // host copy (tau pairing for the rhs outputs)
  // Dirichlet faces (EtoB == 1): slot map [K*4] (-1: none), prescribed states
  // [nvar][nd][Nfp] and the faces of the slots in ascending order.
  int nd = 0;
  int *dmap = nullptr;
  double *gbnd = nullptr;
  bool g_set = false;
  std::vector<int> h_dface;
  // Optional per-kernel timing with HIP events on the context's stream.
  bool timing = false;
  std::vector<hipEvent_t> ev_pool;
  std::vector<int> ev_class;  // kernel class of each recorded (start, stop) pair
  double t_ms[DGLM_NTIMERS] = {0};
  long t_cnt[DGLM_NTIMERS] = {0};
};

extern "C" const char *dglm_last_error(void) { return g_err.c_str(); }
extern "C" int dglm_version(void) { return 100; }

extern "C" int dglm_set_alloc_faults(int n) {
  if (n < 0) return fail(DGLM_ERR_ARG, "dglm_set_alloc_faults: n must be >= 0");
  g_alloc_faults.store(n);
  return DGLM_OK;
}

// ---------------------------------------------------------------------------
// Reference-element tables
// ---------------------------------------------------------------------------
extern "C" int dglm_ref_sizes(int N, int *Np, int *Nfp) {
  if (N < 1 || N > 14) return fail(DGLM_ERR_ARG, "N must be in [1, 14]");
  if (Np) *Np = tet_np(N);
  if (Nfp) *Nfp = tri_np(N);
  return DGLM_OK;
}

extern "C" int dglm_ref_operators(int N, const dglm_ref_tables *out) {
  if (!out) return fail(DGLM_ERR_ARG, "dglm_ref_operators: null output table");
  try {
    RefOps r;
    reference_operators(N, r);
    auto cp = [](const std::vector<double> &v, double *d) {
      if (d) memcpy(d, v.data(), v.size() * sizeof(double));
    };
    cp(r.xyz, out->xyz);
    cp(r.rst, out->rst);
    cp(r.V, out->V);
    cp(r.M, out->M);
    cp(r.invM, out->invM);
    cp(r.Dr, out->Dr);
    cp(r.Ds, out->Ds);
    cp(r.Dt, out->Dt);
    cp(r.WDr, out->WDr);
    cp(r.WDs, out->WDs);
    cp(r.WDt, out->WDt);
    cp(r.Emat, out->Emat);
    cp(r.LIFT, out->LIFT);
    if (out->Fmask) memcpy(out->Fmask, r.Fmask.data(), r.Fmask.size() * sizeof(int));
    if (out->FmaskP) memcpy(out->FmaskP, r.FmaskP.data(), r.FmaskP.size() * sizeof(int));
  } catch (std::exception &ex) {
    return fail(DGLM_ERR_ARG, ex.what());
  }
  return DGLM_OK;
}

extern "C" int dglm_interp_operators(int N, int high_N, double *high_rst, double *interpM, double *high_M) {
  try {
    std::vector<double> a, b, c;
    interp_operators(N, high_N, a, b, c);
    if (high_rst) memcpy(high_rst, a.data(), a.size() * sizeof(double));
    if (interpM) memcpy(interpM, b.data(), b.size() * sizeof(double));
    if (high_M) memcpy(high_M, c.data(), c.size() * sizeof(double));
  } catch (std::exception &ex) {
    return fail(DGLM_ERR_ARG, ex.what());
  }
  return DGLM_OK;
}

extern "C" int dglm_gll_nodes(int N, double *x) {
  if (N < 1 || N > 30 || !x) return fail(DGLM_ERR_ARG, "dglm_gll_nodes: bad arguments");
  std::vector<long double> g = gll_nodes(N);
  for (int i = 0; i <= N; i++) x[i] = (double)g[i];
  return DGLM_OK;
}

extern "C" int dglm_jacobi(int n, double a, double b, int npts, const double *x, double *P, double *dP) {
  if (n < 0 || npts < 0 || !x) return fail(DGLM_ERR_ARG, "dglm_jacobi: bad arguments");
  for (int i = 0; i < npts; i++) {
    if (P) P[i] = (double)jacobi_value(n, a, b, x[i]);
    if (dP) dP[i] = (double)jacobi_grad(n, a, b, x[i]);
  }
  return DGLM_OK;
}

// ---------------------------------------------------------------------------
// Mesh
// ---------------------------------------------------------------------------
extern "C" int dglm_uniform_mesh(int nx, int ny, int nz, const double *box, const int *phy_tags, double *VXYZ,
                                 int *EtoV, int *EtoE, int *EtoF, int *EtoR, int *EtoB) {
  if (!box || !phy_tags || !VXYZ || !EtoV || !EtoE || !EtoF || !EtoR || !EtoB)
    return fail(DGLM_ERR_ARG, "dglm_uniform_mesh: null argument");
  if (nx < 2 || ny < 2 || nz < 2) return fail(DGLM_ERR_ARG, "dglm_uniform_mesh: need nx, ny, nz >= 2");
  try {
    uniform_vertices(nx, ny, nz, box[0], box[1], box[2], box[3], box[4], box[5], VXYZ);
    uniform_connectivity(nx, ny, nz, phy_tags, EtoV, EtoE, EtoF, EtoR, EtoB);
  } catch (std::exception &ex) {
    return fail(DGLM_ERR_ARG, ex.what());
  }
  return DGLM_OK;
}

extern "C" int dglm_mesh_points(int K, int Np, const int *EtoV, const double *VXYZ, const double *rst, double *PX,
                                double *PY, double *PZ) {
  if (K < 0 || Np <= 0 || !EtoV || !VXYZ || !rst || !PX || !PY || !PZ)
    return fail(DGLM_ERR_ARG, "dglm_mesh_points: bad arguments");
  element_points(K, Np, EtoV, VXYZ, rst, PX, PY, PZ);
  return DGLM_OK;
}

extern "C" int dglm_mesh_metrics(int K, const int *EtoV, const double *VXYZ, double *J, double *metrics) {
  if (K < 0 || !EtoV || !VXYZ || !J || !metrics) return fail(DGLM_ERR_ARG, "dglm_mesh_metrics: bad arguments");
  double *m[9];
  for (int i = 0; i < 9; i++) m[i] = metrics + (size_t)i * K;
  element_metrics(K, EtoV, VXYZ, J, m);
  return DGLM_OK;
}

extern "C" int dglm_mesh_normals(int K, const double *J, const double *metrics, double *nx, double *ny, double *nz,
                                 double *sJ, double *Fscale) {
  if (K < 0 || !J || !metrics || !nx || !ny || !nz || !sJ || !Fscale)
    return fail(DGLM_ERR_ARG, "dglm_mesh_normals: bad arguments");
  const double *m[9];
  for (int i = 0; i < 9; i++) m[i] = metrics + (size_t)i * K;
  face_normals(K, J, (double *const *)m, nx, ny, nz, sJ, Fscale);
  return DGLM_OK;
}

extern "C" int dglm_mesh_vmaps(int K, int N, const int *EtoE, const int *EtoF, const int *EtoR, int *vmapM,
                               int *vmapP, int *vmapF) {
  if (K < 0 || !EtoE || !EtoF || !EtoR || !vmapM || !vmapP || !vmapF)
    return fail(DGLM_ERR_ARG, "dglm_mesh_vmaps: bad arguments");
  try {
    RefOps r;
    reference_operators(N, r);
    face_vmaps(K, r.Np, r.Nfp, EtoE, EtoF, EtoR, r.Fmask.data(), r.FmaskP.data(), vmapM, vmapP, vmapF);
  } catch (std::exception &ex) {
    return fail(DGLM_ERR_ARG, ex.what());
  }
  return DGLM_OK;
}

extern "C" int dglm_mesh_spatial_h(int K, const int *EtoV, const double *VXYZ, double *h) {
  if (K < 0 || !EtoV || !VXYZ || !h) return fail(DGLM_ERR_ARG, "dglm_mesh_spatial_h: bad arguments");
  spatial_h(K, EtoV, VXYZ, h);
  return DGLM_OK;
}

// ---------------------------------------------------------------------------
// Kernel dispatch
// ---------------------------------------------------------------------------
namespace {

// Wave speed of state u into maxv, with the per-block maxima of lambda*Fscale
// (and their face indices) in partial / partial_idx [nblk_ws] for the dt
// kernel.
template <int N, int PHYS>
void launch_wavespeed(dglm_ctx *c, const double *u, double *maxv, double *partial, int *partial_idx) {
  hipLaunchKernelGGL((wavespeed_lds_kernel<N, PHYS>), dim3(c->nblk_ws), dim3(256), 0, c->stream, c->K, u, c->vel,
                     c->geo, c->fmx, maxv, partial, partial_idx, c->p.gamma);
}

// True when the MFMA stage kernel is built for this order.
inline bool mfma_available(int N) { return N >= 1 && N <= 6; }

// True when the warp-specialised kernel exists for this order.
inline bool ws_built(int N) {
  switch (N) {
    case 3:
      return ws_available<3>();
    case 4:
      return ws_available<4>();
    default:
      return false;
  }
}

// The stage kernel that runs: the request where it exists, else the best
// available: the warp-specialised kernel (FP64, and DGLM_FP32 at N = 3, 4),
// else the fused mixed-precision kernel (DGLM_FP32 Euler, N <= 6), else the
// generic MFMA kernel, else the LDS-staged one.
int stage_kind(const dglm_ctx *c) {
  const bool ws = c->afrag_ws != nullptr, mf = mfma_available(c->p.N) && c->afrag != nullptr;
  if (c->variant == DGLM_STAGE_SIMPLE && c->p.precision == DGLM_FP64) return DGLM_STAGE_SIMPLE;
  if (c->variant == DGLM_STAGE_MFMA && mf) return DGLM_STAGE_MFMA;
  if (c->variant == DGLM_STAGE_MFMA_MX && c->mx_tab) return DGLM_STAGE_MFMA_MX;
  if (ws) return DGLM_STAGE_MFMA_WS;
  if (c->mx_tab) return DGLM_STAGE_MFMA_MX;
  return mf ? DGLM_STAGE_MFMA : DGLM_STAGE_SIMPLE;
}

bool use_mfma(const dglm_ctx *c) { return stage_kind(c) == DGLM_STAGE_MFMA; }

template <int N, int MODE>
void launch_ws(dglm_ctx *c, const StageArgs &a) {
  if constexpr (ws_available<N>()) {
    using S = WsShape<N>;
    const int nbatch = (c->K + S::EB - 1) / S::EB;
    if (c->ws_grid == 0) {
      int dev = 0, cus = 0, occ = 0;
      hipGetDevice(&dev);
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
      hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, stage_ws_kernel<N, 0, false>, S::NT, 0);
      c->ws_grid = (cus > 0 ? cus : 256) * (occ > 0 ? occ : 1);
      c->ws_groups = c->ws_grid % 8 == 0 ? 8 : 1;
    }
    int grid = c->ws_grid, groups = c->ws_groups;
    if (nbatch < grid) {
      grid = nbatch;
      groups = 1;
    }
    if (a.dmap)
      launch_ws_dirichlet(N, MODE, grid, c->stream, a, c->afrag_ws, c->arem_ws, c->ws_order, nbatch, groups);
    else
      hipLaunchKernelGGL((stage_ws_kernel<N, MODE, false>), dim3(grid), dim3(S::NT), 0, c->stream, a, c->afrag_ws,
                         c->arem_ws, c->ws_order, nbatch, groups);
  }
}

template <int N, int MODE>
void launch_mx(dglm_ctx *c, const StageArgs &a) {
  if constexpr (N <= 6) {
    using S = MxShape<N>;
    const int nbatch = (c->K + S::EB - 1) / S::EB;
    if (c->mx_grid == 0) {
      int dev = 0, cus = 0, occ = 0;
      hipGetDevice(&dev);
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
      hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, stage_mx_kernel<N, 0>, S::NT, 0);
      c->mx_grid = (cus > 0 ? cus : 256) * (occ > 0 ? occ : 1);
      c->mx_groups = c->mx_grid % 8 == 0 ? 8 : 1;
    }
    int grid = c->mx_grid, groups = c->mx_groups;
    if (nbatch < grid) {
      grid = nbatch;
      groups = 1;
    }
    hipLaunchKernelGGL((stage_mx_kernel<N, MODE>), dim3(grid), dim3(S::NT), 0, c->stream, a, c->mx_tab, c->ws_order,
                       nbatch, groups);
  }
}

template <int N, int PHYS, int MODE, int PREC>
void launch_mfma(dglm_ctx *c, const StageArgs &a) {
  using S = MfmaShape<N>;
  const int nbatch = (c->K + 15) / 16;
  if (c->mfma_grid == 0) {
    int dev = 0, cus = 0, occ = 0;
    hipGetDevice(&dev);
    hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, stage_mfma_kernel<N, PHYS, 0, PREC>, S::NT, 0);
    c->mfma_grid = (cus > 0 ? cus : 256) * (occ > 0 ? occ : 1);
  }
  const int grid = nbatch < c->mfma_grid ? nbatch : c->mfma_grid;
  hipLaunchKernelGGL((stage_mfma_kernel<N, PHYS, MODE, PREC>), dim3(grid), dim3(S::NT), 0, c->stream, a,
                     (const void *)c->afrag, c->rowsum, nbatch);
}

template <int N, int PHYS, int MODE>
void launch_stage(dglm_ctx *c, const StageArgs &a) {
  const int kind = stage_kind(c);
  if (PHYS == PHYS_EULER && kind == DGLM_STAGE_MFMA_WS) {
    launch_ws<N, MODE>(c, a);
    return;
  }
  if (PHYS == PHYS_EULER && kind == DGLM_STAGE_MFMA_MX) {
    launch_mx<N, MODE>(c, a);
    return;
  }
  if constexpr (N <= 6) {
    if (kind == DGLM_STAGE_MFMA) {
      if (c->p.precision == DGLM_FP32)
        launch_mfma<N, PHYS, MODE, PREC_F32>(c, a);
      else if (c->p.precision == DGLM_BF16X3 || c->p.precision == DGLM_BF16X6)
        launch_mfma<N, PHYS, MODE, PREC_BF>(c, a);
      else
        launch_mfma<N, PHYS, MODE, PREC_F64>(c, a);
      return;
    }
  }
  constexpr int EB = eb_simple<N, PHYS>();
  int nblk = (c->K + EB - 1) / EB;
  hipLaunchKernelGGL((stage_simple_kernel<N, PHYS, MODE>), dim3(nblk), dim3(256), 0, c->stream, a);
}

typedef void (*ws_fn)(dglm_ctx *, const double *, double *, double *, int *);
typedef void (*st_fn)(dglm_ctx *, const StageArgs &);

template <int PHYS>
bool pick(int N, ws_fn &ws, st_fn &st0, st_fn &st1) {
  switch (N) {
#define DGLM_CASE(n)                                   \
  case n:                                              \
    ws = launch_wavespeed<n, PHYS>;                    \
    st0 = launch_stage<n, PHYS, 0>;                    \
    st1 = launch_stage<n, PHYS, 1>;                    \
    return true;
    DGLM_CASE(1)
    DGLM_CASE(2)
    DGLM_CASE(3)
    DGLM_CASE(4)
    DGLM_CASE(5)
    DGLM_CASE(6)
    DGLM_CASE(7)
    DGLM_CASE(8)
    DGLM_CASE(9)
    DGLM_CASE(10)
#undef DGLM_CASE
    default:
      return false;
  }
}

bool dispatch(const dglm_ctx *c, ws_fn &ws, st_fn &st0, st_fn &st1) {
  return c->p.physics == DGLM_EULER ? pick<PHYS_EULER>(c->p.N, ws, st0, st1)
                                    : pick<PHYS_ADV>(c->p.N, ws, st0, st1);
}

StageArgs stage_args(dglm_ctx *c, int s) {
  StageArgs a{};
  a.K = c->K;
  a.u_in = c->u[c->cur];
  a.u_out = c->u[c->cur ^ 1];
  a.k = c->k;
  a.rhs = c->rhs;
  a.vel = c->vel;
  a.geo = c->geo;
  a.conn = c->conn;
  a.fmx = c->fmx;
  a.op = c->op;
  a.maxv = c->maxv;
  a.scal = c->scal;
  a.rka = s >= 0 ? RKA[s] : 0.0;
  a.rkb = s >= 0 ? RKB[s] : 0.0;
  a.gamma = c->p.gamma;
  a.tau_scale = c->p.tau_scale;
  a.flux = c->p.flux == DGLM_FLUX_RUSANOV ? 1 : 0;
  a.bf_terms = c->p.precision == DGLM_BF16X6 ? 6 : 3;
  a.dmap = c->nd > 0 ? c->dmap : nullptr;
  a.gbnd = c->gbnd;
  a.nd = c->nd;
  return a;
}

void free_tape(struct dglm_adj_tape *T);

hipError_t device_malloc(void **p, size_t bytes) {
  int n = g_alloc_faults.load();
  while (n > 0)
    if (g_alloc_faults.compare_exchange_weak(n, n - 1)) {
      *p = nullptr;
      return hipErrorOutOfMemory;
    }
  return hipMalloc(p, bytes);
}

template <typename T>
int dev_alloc(T **p, size_t n, const char *what) {
  hipError_t e = device_malloc((void **)p, n * sizeof(T) + 16);
  if (e != hipSuccess) return fail(DGLM_ERR_HIP, std::string("hipMalloc(") + what + "): " + hipGetErrorString(e));
  return DGLM_OK;
}

void free_ctx(dglm_ctx *c) {
  if (!c) return;
  hipSetDevice(c->dev);
  if (c->stream) hipStreamSynchronize(c->stream);
  void *ptrs[] = {c->u[0], c->u[1], c->k, c->rhs, c->vel, c->geo, c->op, c->maxv, c->partial, c->dt_hist,
                  c->conn, c->fmx, c->scal, c->afrag, c->rowsum, c->partial_idx, c->afrag_ws, c->arem_ws,
                  c->ws_order, c->dmap, c->gbnd, c->mx_tab};
  for (void *q : ptrs)
    if (q) hipFree(q);
  free_tape(c->tape);
  for (int g = 0; g < 2; g++)
    if (c->graph[g]) hipGraphExecDestroy(c->graph[g]);
  if (c->d_J) hipFree(c->d_J);
  if (c->d_w) hipFree(c->d_w);
  if (c->d_ipart) hipFree(c->d_ipart);
  for (hipEvent_t e : c->ev_pool) hipEventDestroy(e);
  if (c->stream) hipStreamDestroy(c->stream);
  delete c;
}

// Record the start event of a timed launch; returns the pair index or -1.
int timer_start(dglm_ctx *c, int cls) {
  if (!c->timing) return -1;
  size_t idx = c->ev_class.size();
  while (c->ev_pool.size() < 2 * (idx + 1)) {
    hipEvent_t e;
    if (hipEventCreate(&e) != hipSuccess) return -1;
    c->ev_pool.push_back(e);
  }
  c->ev_class.push_back(cls);
  hipEventRecord(c->ev_pool[2 * idx], c->stream);
  return (int)idx;
}

void timer_stop(dglm_ctx *c, int idx) {
  if (idx >= 0) hipEventRecord(c->ev_pool[2 * idx + 1], c->stream);
}

int timer_flush(dglm_ctx *c) {
  if (c->ev_class.empty()) return DGLM_OK;
  hipError_t e = hipStreamSynchronize(c->stream);
  if (e != hipSuccess) return fail(DGLM_ERR_HIP, std::string("hipStreamSynchronize: ") + hipGetErrorString(e));
  for (size_t i = 0; i < c->ev_class.size(); i++) {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, c->ev_pool[2 * i], c->ev_pool[2 * i + 1]) == hipSuccess) {
      c->t_ms[c->ev_class[i]] += ms;
      c->t_cnt[c->ev_class[i]] += 1;
    }
  }
  c->ev_class.clear();
  return DGLM_OK;
}

}  // namespace

// ---------------------------------------------------------------------------
// Allocation-free stepping (dglm_step_noalloc; no reference counterpart):
// its device scratch -- the dt history and the wave-speed block maxima with
// their face indices -- lives in this fixed array instead of buffers the
// library allocates.  Each device holds its own copy until the process ends;
// every call on one device shares it, so calls that overlap on one device
// race.  Layout in doubles: [0, MAX_STEPS) dt history, [MAX_STEPS, +MAX_BLOCKS)
// block maxima, then MAX_BLOCKS int32 face indices.
// ---------------------------------------------------------------------------
constexpr size_t NOALLOC_DOUBLES = (size_t)DGLM_NOALLOC_MAX_STEPS + DGLM_NOALLOC_MAX_BLOCKS + DGLM_NOALLOC_MAX_BLOCKS / 2;
static_assert(NOALLOC_DOUBLES * sizeof(double) == DGLM_NOALLOC_SCRATCH_BYTES, "scratch size");
static __device__ __align__(256) double scratch[NOALLOC_DOUBLES];

namespace {
std::mutex g_scratch_mu;
std::vector<double *> g_scratch_ptr;  // device address of `scratch` per device ordinal (null: not resolved)

// The device address of `scratch` on device dev, resolved once per device
// with that device current (the array's name is only a host-side stand-in).
int scratch_on(int dev, double **out) {
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (dev < 0) return fail(DGLM_ERR_ARG, "scratch: bad device ordinal");
  if ((int)g_scratch_ptr.size() <= dev) g_scratch_ptr.resize(dev + 1, nullptr);
  if (!g_scratch_ptr[dev]) {
    void *p = nullptr;
    HIP_TRY(hipSetDevice(dev));
    HIP_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) return fail(DGLM_ERR_HIP, "hipGetSymbolAddress(scratch): null address");
    g_scratch_ptr[dev] = (double *)p;
  }
  *out = g_scratch_ptr[dev];
  return DGLM_OK;
}
}  // namespace

// ---------------------------------------------------------------------------
// Context
// ---------------------------------------------------------------------------
extern "C" int dglm_create(const dglm_params *p, const dglm_mesh *m, dglm_ctx **out) {
  if (!p || !m || !out) return fail(DGLM_ERR_ARG, "dglm_create: null argument");
  *out = nullptr;
  if (p->physics != DGLM_ADVECTION && p->physics != DGLM_EULER)
    return fail(DGLM_ERR_ARG, "dglm_create: physics must be DGLM_ADVECTION or DGLM_EULER");
  if (p->precision != DGLM_FP64 && p->precision != DGLM_FP32 && p->precision != DGLM_BF16X3 &&
      p->precision != DGLM_BF16X6)
    return fail(DGLM_ERR_ARG, "dglm_create: unknown precision mode");
  if (p->precision != DGLM_FP64 && (p->N < 1 || p->N > 6))
    return fail(DGLM_ERR_UNSUPPORTED, "dglm_create: mixed precision is built for N in [1, 6]");
  if (p->flux != DGLM_FLUX_LM && p->flux != DGLM_FLUX_RUSANOV)
    return fail(DGLM_ERR_ARG, "dglm_create: flux must be DGLM_FLUX_LM or DGLM_FLUX_RUSANOV");
  if (p->N < 1 || p->N > 10) return fail(DGLM_ERR_UNSUPPORTED, "dglm_create: N must be in [1, 10]");
  if (m->K <= 0) return fail(DGLM_ERR_ARG, "dglm_create: K must be positive");
  if (!(p->cfl > 0) || !(p->tau_scale >= 0) || !(p->gamma > 1.0) || std::isnan(p->tmax))
    return fail(DGLM_ERR_ARG, "dglm_create: invalid cfl / gamma / tau_scale / tmax");
  const double *req[] = {m->WDr, m->WDs, m->WDt, m->LIFT, m->Fscale, m->rx, m->ry, m->rz, m->sx,
                         m->sy,  m->sz,  m->tx,  m->ty,   m->tz,     m->nx, m->ny, m->nz};
  for (const double *q : req)
    if (!q) return fail(DGLM_ERR_ARG, "dglm_create: missing mesh / operator array");
  if (!m->vmapP || !m->vmapF || !m->Fmask) return fail(DGLM_ERR_ARG, "dglm_create: missing index array");

  RefOps ref;
  try {
    reference_operators(p->N, ref);
  } catch (std::exception &ex) {
    return fail(DGLM_ERR_ARG, ex.what());
  }
  const int K = m->K, Np = ref.Np, Nfp = ref.Nfp, NFF = 4 * Nfp, KD = 3 * Np + NFF;
  for (int i = 0; i < 4 * Nfp; i++)
    if (m->Fmask[i] != ref.Fmask[i]) return fail(DGLM_ERR_ARG, "dglm_create: Fmask does not match order N");

  // Face-node table: code 0..2 = FmaskP[f][rot], code 3 = Fmask[f].
  std::vector<int> fmx(16 * Nfp);
  for (int f = 0; f < 4; f++)
    for (int cde = 0; cde < 4; cde++)
      for (int j = 0; j < Nfp; j++)
        fmx[(f * 4 + cde) * Nfp + j] = cde < 3 ? ref.FmaskP[(f * 3 + cde) * Nfp + j] : ref.Fmask[f * Nfp + j];

  // Packed connectivity from (vmapF, vmapP): find the neighbour face and the
  // node permutation that reproduces vmapP exactly.
  std::vector<int> conn, dmap, dface;
  {
    std::string err;
    if (int rc = pack_connectivity(K, Np, Nfp, m->vmapP, m->vmapF, m->EtoB, fmx.data(), conn, dmap, dface, err))
      return fail(rc, "dglm_create: " + err);
  }

  // Geometry records.
  std::vector<double> geo((size_t)K * GEO_STRIDE);
  const double *mt[9] = {m->rx, m->ry, m->rz, m->sx, m->sy, m->sz, m->tx, m->ty, m->tz};
  for (int e = 0; e < K; e++) {
    double *g = &geo[(size_t)e * GEO_STRIDE];
    for (int q = 0; q < 9; q++) g[q] = mt[q][e];
    for (int f = 0; f < 4; f++) {
      g[9 + f * 4 + 0] = m->nx[e * 4 + f];
      g[9 + f * 4 + 1] = m->ny[e * 4 + f];
      g[9 + f * 4 + 2] = m->nz[e * 4 + f];
      g[9 + f * 4 + 3] = m->Fscale[e * 4 + f];
    }
  }

  // Stacked operator, column-major [KD][Np]: op[k][i].
  std::vector<double> op((size_t)KD * Np);
  auto at = [&](const double *A, int rows, int cols, int i, int j) {
    return m->matrix_order == 0 ? A[(size_t)i * cols + j] : A[(size_t)j * rows + i];
  };
  const double *WD[3] = {m->WDr, m->WDs, m->WDt};
  for (int d = 0; d < 3; d++)
    for (int j = 0; j < Np; j++)
      for (int i = 0; i < Np; i++) op[(size_t)(d * Np + j) * Np + i] = at(WD[d], Np, Np, i, j);
  for (int q = 0; q < NFF; q++)
    for (int i = 0; i < Np; i++) op[(size_t)(3 * Np + q) * Np + i] = -at(m->LIFT, Np, NFF, i, q);

  // A-fragments of the stacked operator for the MFMA stage kernel.
  std::vector<double> afr;
  std::vector<float> afr32;
  std::vector<s16x4> afrbf;
  std::vector<double> rsum;
  if (mfma_available(p->N)) {
    const int RT = (Np + 15) / 16, KS = (KD + 3) / 4;
    if (p->precision != DGLM_FP64) {
      if (p->precision == DGLM_FP32) {
        afr32.resize((size_t)RT * KS * 64);
        arrange_a_fragments(Np, KD, RT, KS, op.data(), afr32.data());
      } else {
        const int KS16 = (KD + 15) / 16;
        afrbf.resize((size_t)RT * KS16 * BF_PIECES * 64);
        arrange_a_bf16(Np, KD, RT, KS16, op.data(), afrbf.data());
      }
      // affine-offset tables: [fitV (KD x 4) | fitP (KD x 4) | fitAV (Np x NBLK x 4)]
      rsum.assign((size_t)8 * KD + (size_t)Np * NBLK * 4, 0.0);
      std::vector<double> rx(Np), ry(Np), rz(Np);
      for (int i = 0; i < Np; i++) {
        rx[i] = ref.rst[i * 3];
        ry[i] = ref.rst[i * 3 + 1];
        rz[i] = ref.rst[i * 3 + 2];
      }
      if (!affine_fit_tables(Np, Nfp, ref.Fmask.data(), rx.data(), ry.data(), rz.data(), op.data(), rsum.data(),
                             rsum.data() + 4 * KD, rsum.data() + 8 * KD))
        return fail(DGLM_ERR_ARG, "dglm_create: singular affine-offset basis");
    } else {
      afr.resize((size_t)RT * KS * 64);
      arrange_a_fragments(Np, KD, RT, KS, op.data(), afr.data());
    }
  }
  const size_t afrag_bytes = afr.size() * sizeof(double) + afr32.size() * sizeof(float) + afrbf.size() * sizeof(s16x4);
  const void *afrag_src = !afr.empty() ? (const void *)afr.data()
                                       : (!afr32.empty() ? (const void *)afr32.data() : (const void *)afrbf.data());
  // Operator rows for the warp-specialised FP64 Euler kernel, which also runs
  // DGLM_FP32 at the orders it is built for (N = 3, 4): there its FP64
  // contraction outruns the fused mixed-precision kernel, whose least-squares
  // fits cost more FP64 work than the FP32 MFMA saves when the stage is bound
  // by the FP64 flux arithmetic, and it is exact where FP32 promises 1e-8.
  std::vector<double> afr_ws, arem_ws;
  // (its gathers index one variable's nodes with 32-bit offsets)
  if (p->physics == DGLM_EULER && (p->precision == DGLM_FP64 || p->precision == DGLM_FP32) && ws_built(p->N) &&
      (long long)K * Np < (1LL << 31)) {
    if (p->N == 3) {
      afr_ws.resize((size_t)WsShape<3>::RTF * WsShape<3>::KS * 64);
      arem_ws.resize((size_t)WsShape<3>::REM * WsShape<3>::KDP);
      ws_arrange<3>(op.data(), afr_ws.data(), arem_ws.data());
    } else {
      afr_ws.resize((size_t)WsShape<4>::RTF * WsShape<4>::KS * 64);
      arem_ws.resize((size_t)WsShape<4>::REM * WsShape<4>::KDP);
      ws_arrange<4>(op.data(), afr_ws.data(), arem_ws.data());
    }
  }

  // Tables of the fused mixed-precision Euler kernel (no Dirichlet faces:
  // those take the generic MFMA kernel).
  std::vector<unsigned char> mx_tab;
  if (p->physics == DGLM_EULER && p->precision == DGLM_FP32 && dface.empty() && !rsum.empty() &&
      (long long)K * Np < (1LL << 31)) {
    const double *fitV = rsum.data(), *fitP = rsum.data() + 4 * (size_t)KD, *fitAV = rsum.data() + 8 * (size_t)KD;
    switch (p->N) {
#define DGLM_MX_CASE(n)                                                  \
  case n:                                                              \
    mx_tab.assign(MxTables<n>::bytes, 0);                              \
    mx_build_tables<n>(op.data(), fitV, fitP, fitAV, mx_tab.data());   \
    break;
      DGLM_MX_CASE(1)
      DGLM_MX_CASE(2)
      DGLM_MX_CASE(3)
      DGLM_MX_CASE(4)
      DGLM_MX_CASE(5)
      DGLM_MX_CASE(6)
#undef DGLM_MX_CASE
      default:
        break;
    }
  }

  // Batch order of the warp-specialised kernel: 16-element batches sorted by
  // the Morton code of their node centroid, so that the workgroups of one XCD,
  // which walk a contiguous stretch of the order together, work on a compact
  // region and find the neighbours' face traces in their L2.
  std::vector<int> ws_order;
  if ((!afr_ws.empty() || !mx_tab.empty()) && m->px && m->py && m->pz) {
    const int nb = (K + 15) / 16;
    const size_t KNp = (size_t)K * Np;
    double lo[3] = {HUGE_VAL, HUGE_VAL, HUGE_VAL}, hi[3] = {-HUGE_VAL, -HUGE_VAL, -HUGE_VAL};
    const double *pc[3] = {m->px, m->py, m->pz};
    std::vector<double> cen((size_t)nb * 3, 0.0);
    for (int d = 0; d < 3; d++)
      for (size_t i = 0; i < KNp; i++) {
        const double x = pc[d][i];
        lo[d] = x < lo[d] ? x : lo[d];
        hi[d] = x > hi[d] ? x : hi[d];
        cen[(i / Np / 16) * 3 + d] += x;
      }
    std::vector<std::pair<unsigned long long, int>> key(nb);
    for (int bi = 0; bi < nb; bi++) {
      const int ne = std::min(16, K - 16 * bi);
      unsigned long long code = 0;
      unsigned q[3];
      for (int d = 0; d < 3; d++) {
        const double span = hi[d] > lo[d] ? hi[d] - lo[d] : 1.0;
        const double t = (cen[(size_t)bi * 3 + d] / ((double)ne * Np) - lo[d]) / span;
        q[d] = (unsigned)std::min(1023.0, std::max(0.0, t * 1024.0));
      }
      for (int bit = 9; bit >= 0; bit--)
        for (int d = 0; d < 3; d++) code = (code << 1) | ((q[d] >> bit) & 1u);
      key[bi] = {code, bi};
    }
    std::stable_sort(key.begin(), key.end());
    ws_order.resize(nb);
    for (int bi = 0; bi < nb; bi++) ws_order[bi] = key[bi].second;
  }

  dglm_ctx *c = new dglm_ctx();
  c->p = *p;
  c->K = K;
  c->Np = Np;
  c->Nfp = Nfp;
  c->NFF = NFF;
  c->KD = KD;
  c->nvar = p->physics == DGLM_EULER ? 5 : 1;
  c->dev = p->device;
  c->nblk_ws = (K + wavespeed_elems(Np) - 1) / wavespeed_elems(Np);
  c->h_conn = conn;
  c->nd = (int)dface.size();
  c->h_dface = dface;
  int rc;
  hipError_t he = hipSetDevice(c->dev);
  if (he != hipSuccess) {
    delete c;
    return fail(DGLM_ERR_HIP, std::string("hipSetDevice: ") + hipGetErrorString(he));
  }
  if ((rc = scratch_on(c->dev, &c->scratch))) {
    delete c;
    return rc;
  }
  const size_t nst = (size_t)c->nvar * K * Np;
  if ((rc = dev_alloc(&c->u[0], nst, "u0")) || (rc = dev_alloc(&c->u[1], nst, "u1")) ||
      (rc = dev_alloc(&c->k, nst, "k")) || (rc = dev_alloc(&c->geo, geo.size(), "geo")) ||
      (rc = dev_alloc(&c->op, op.size(), "op")) || (rc = dev_alloc(&c->maxv, (size_t)K * 4, "maxv")) ||
      (rc = dev_alloc(&c->partial, (size_t)c->nblk_ws, "partial")) ||
      (rc = dev_alloc(&c->partial_idx, (size_t)c->nblk_ws, "partial index")) ||
      (rc = dev_alloc(&c->conn, conn.size(), "conn")) || (rc = dev_alloc(&c->fmx, fmx.size(), "fmx")) ||
      (rc = dev_alloc(&c->scal, 1, "scalars")) ||
      (afrag_bytes && (rc = dev_alloc((char **)&c->afrag, afrag_bytes, "A-fragments"))) ||
      (!afr_ws.empty() && (rc = dev_alloc(&c->afrag_ws, afr_ws.size(), "A-fragments (warp-specialised)"))) ||
      (!arem_ws.empty() && (rc = dev_alloc(&c->arem_ws, arem_ws.size(), "remainder rows"))) ||
      (!ws_order.empty() && (rc = dev_alloc(&c->ws_order, ws_order.size(), "batch order"))) ||
      (!rsum.empty() && (rc = dev_alloc(&c->rowsum, rsum.size(), "affine-offset tables"))) ||
      (!mx_tab.empty() && (rc = dev_alloc(&c->mx_tab, mx_tab.size(), "mixed-precision tables")))) {
    free_ctx(c);
    return rc;
  }
  if (c->nvar == 1 && (rc = dev_alloc(&c->vel, (size_t)3 * K * Np, "velocity"))) {
    free_ctx(c);
    return rc;
  }
  if (c->nd > 0 && ((rc = dev_alloc(&c->dmap, dmap.size(), "Dirichlet slot map")) ||
                    (rc = dev_alloc(&c->gbnd, (size_t)c->nvar * c->nd * Nfp, "Dirichlet states")))) {
    free_ctx(c);
    return rc;
  }
  he = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
  if (he != hipSuccess) {
    free_ctx(c);
    return fail(DGLM_ERR_HIP, std::string("hipStreamCreate: ") + hipGetErrorString(he));
  }
  DevScalars s0{};
  s0.active = 1;
  hipError_t errs[] = {
      hipMemcpy(c->geo, geo.data(), geo.size() * sizeof(double), hipMemcpyHostToDevice),
      hipMemcpy(c->op, op.data(), op.size() * sizeof(double), hipMemcpyHostToDevice),
      hipMemcpy(c->conn, conn.data(), conn.size() * sizeof(int), hipMemcpyHostToDevice),
      hipMemcpy(c->fmx, fmx.data(), fmx.size() * sizeof(int), hipMemcpyHostToDevice),
      c->afrag ? hipMemcpy(c->afrag, afrag_src, afrag_bytes, hipMemcpyHostToDevice) : hipSuccess,
      c->afrag_ws ? hipMemcpy(c->afrag_ws, afr_ws.data(), afr_ws.size() * sizeof(double), hipMemcpyHostToDevice)
                  : hipSuccess,
      c->arem_ws ? hipMemcpy(c->arem_ws, arem_ws.data(), arem_ws.size() * sizeof(double), hipMemcpyHostToDevice)
                 : hipSuccess,
      c->ws_order ? hipMemcpy(c->ws_order, ws_order.data(), ws_order.size() * sizeof(int), hipMemcpyHostToDevice)
                  : hipSuccess,
      c->rowsum ? hipMemcpy(c->rowsum, rsum.data(), rsum.size() * sizeof(double), hipMemcpyHostToDevice)
                : hipSuccess,
      c->dmap ? hipMemcpy(c->dmap, dmap.data(), dmap.size() * sizeof(int), hipMemcpyHostToDevice) : hipSuccess,
      c->mx_tab ? hipMemcpy(c->mx_tab, mx_tab.data(), mx_tab.size(), hipMemcpyHostToDevice) : hipSuccess,
      hipMemcpy(c->scal, &s0, sizeof(s0), hipMemcpyHostToDevice),
      hipMemset(c->u[0], 0, nst * sizeof(double)),
      hipMemset(c->u[1], 0, nst * sizeof(double)),
      hipMemset(c->k, 0, nst * sizeof(double)),
  };
  for (hipError_t e : errs)
    if (e != hipSuccess) {
      free_ctx(c);
      return fail(DGLM_ERR_HIP, std::string("upload: ") + hipGetErrorString(e));
    }
  if (c->vel) {
    // Default advection velocity (1,1,1), as advec_init.cpp:17-19.
    std::vector<double> one((size_t)3 * K * Np, 1.0);
    if (hipMemcpy(c->vel, one.data(), one.size() * sizeof(double), hipMemcpyHostToDevice) != hipSuccess) {
      free_ctx(c);
      return fail(DGLM_ERR_HIP, "upload velocity failed");
    }
  }
  *out = c;
  return DGLM_OK;
}

extern "C" int dglm_create_from_dir(const dglm_params *p, const char *datain_dir, int matrix_order, dglm_ctx **out) {
  if (!p || !datain_dir || !out) return fail(DGLM_ERR_ARG, "dglm_create_from_dir: null argument");
  if (p->N < 1 || p->N > 14) return fail(DGLM_ERR_ARG, "dglm_create_from_dir: bad N");
  Datain d;
  std::string err;
  if (int rc = read_datain(datain_dir, tet_np(p->N), tri_np(p->N), d, err)) return fail(rc, err);
  dglm_mesh m{};
  m.K = d.K;
  m.matrix_order = matrix_order;
  m.EtoB = d.EtoB.data();
  m.vmapP = d.vmapP.data();
  m.vmapF = d.vmapF.data();
  m.Fmask = d.Fmask.data();
  m.WDr = d.WDr.data();
  m.WDs = d.WDs.data();
  m.WDt = d.WDt.data();
  m.LIFT = d.LIFT.data();
  m.Fscale = d.Fscale.data();
  const double **mt[9] = {&m.rx, &m.ry, &m.rz, &m.sx, &m.sy, &m.sz, &m.tx, &m.ty, &m.tz};
  for (int q = 0; q < 9; q++) *mt[q] = d.metric[q].data();
  m.nx = d.nx.data();
  m.ny = d.ny.data();
  m.nz = d.nz.data();
  m.px = d.px.data();
  m.py = d.py.data();
  m.pz = d.pz.data();
  return dglm_create(p, &m, out);
}

extern "C" int dglm_destroy(dglm_ctx *c) {
  free_ctx(c);
  return DGLM_OK;
}

extern "C" int dglm_sizes(const dglm_ctx *c, int *K, int *Np, int *Nfp, int *nvar) {
  if (!c) return fail(DGLM_ERR_ARG, "dglm_sizes: null context");
  if (K) *K = c->K;
  if (Np) *Np = c->Np;
  if (Nfp) *Nfp = c->Nfp;
  if (nvar) *nvar = c->nvar;
  return DGLM_OK;
}

static int set_state_impl(dglm_ctx *c, const double *u, hipMemcpyKind kind) {
  if (!c || !u) return fail(DGLM_ERR_ARG, "dglm_set_state: null argument");
  HIP_TRY(hipSetDevice(c->dev));
  const size_t n = (size_t)c->nvar * c->K * c->Np;
  c->cur = 0;
  HIP_TRY(hipMemcpyAsync(c->u[0], u, n * sizeof(double), kind, c->stream));
  HIP_TRY(hipMemsetAsync(c->k, 0, n * sizeof(double), c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  return DGLM_OK;
}

extern "C" int dglm_set_state(dglm_ctx *c, const double *u) { return set_state_impl(c, u, hipMemcpyHostToDevice); }
extern "C" int dglm_set_state_device(dglm_ctx *c, const double *d_u) {
  return set_state_impl(c, d_u, hipMemcpyDeviceToDevice);
}

static int get_state_impl(dglm_ctx *c, double *u, hipMemcpyKind kind) {
  if (!c || !u) return fail(DGLM_ERR_ARG, "dglm_get_state: null argument");
  HIP_TRY(hipSetDevice(c->dev));
  const size_t n = (size_t)c->nvar * c->K * c->Np;
  HIP_TRY(hipMemcpyAsync(u, c->u[c->cur], n * sizeof(double), kind, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  return DGLM_OK;
}

extern "C" int dglm_get_state(dglm_ctx *c, double *u) { return get_state_impl(c, u, hipMemcpyDeviceToHost); }
extern "C" int dglm_get_state_device(dglm_ctx *c, double *d_u) {
  return get_state_impl(c, d_u, hipMemcpyDeviceToDevice);
}

extern "C" int dglm_set_velocity(dglm_ctx *c, const double *vx, const double *vy, const double *vz) {
  if (!c || !vx || !vy || !vz) return fail(DGLM_ERR_ARG, "dglm_set_velocity: null argument");
  if (!c->vel) return fail(DGLM_ERR_ARG, "dglm_set_velocity: only for advection");
  HIP_TRY(hipSetDevice(c->dev));
  const size_t n = (size_t)c->K * c->Np;
  HIP_TRY(hipMemcpy(c->vel, vx, n * sizeof(double), hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(c->vel + n, vy, n * sizeof(double), hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(c->vel + 2 * n, vz, n * sizeof(double), hipMemcpyHostToDevice));
  return DGLM_OK;
}

extern "C" int dglm_dirichlet_faces(const dglm_ctx *c, int *nd, int *faces) {
  if (!c) return fail(DGLM_ERR_ARG, "dglm_dirichlet_faces: null context");
  if (nd) *nd = c->nd;
  if (faces && c->nd > 0) memcpy(faces, c->h_dface.data(), c->h_dface.size() * sizeof(int));
  return DGLM_OK;
}

static int set_dirichlet_impl(dglm_ctx *c, const double *g, hipMemcpyKind kind) {
  if (!c) return fail(DGLM_ERR_ARG, "dglm_set_dirichlet: null context");
  if (c->nd == 0) return fail(DGLM_ERR_ARG, "dglm_set_dirichlet: the mesh has no Dirichlet faces (EtoB == 1)");
  if (!g) return fail(DGLM_ERR_ARG, "dglm_set_dirichlet: null states");
  HIP_TRY(hipSetDevice(c->dev));
  const size_t n = (size_t)c->nvar * c->nd * c->Nfp;
  HIP_TRY(hipMemcpyAsync(c->gbnd, g, n * sizeof(double), kind, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  c->g_set = true;
  return DGLM_OK;
}

extern "C" int dglm_set_dirichlet(dglm_ctx *c, const double *g) {
  return set_dirichlet_impl(c, g, hipMemcpyHostToDevice);
}
extern "C" int dglm_set_dirichlet_device(dglm_ctx *c, const double *d_g) {
  return set_dirichlet_impl(c, d_g, hipMemcpyDeviceToDevice);
}

// Stepping a mesh with Dirichlet faces needs their states.
static int need_dirichlet(const dglm_ctx *c, const char *who) {
  if (c->nd > 0 && !c->g_set)
    return fail(DGLM_ERR_ARG, std::string(who) + ": the mesh has Dirichlet faces (EtoB == 1) and no prescribed "
                                                 "states (dglm_set_dirichlet)");
  return DGLM_OK;
}

extern "C" int dglm_set_time(dglm_ctx *c, double t, int tstep) {
  if (!c) return fail(DGLM_ERR_ARG, "dglm_set_time: null context");
  HIP_TRY(hipSetDevice(c->dev));
  DevScalars s{};
  s.t = t;
  s.t_next = t;
  s.tstep = tstep;
  s.active = 1;
  HIP_TRY(hipMemcpyAsync(c->scal, &s, sizeof(s), hipMemcpyHostToDevice, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  return DGLM_OK;
}

extern "C" int dglm_get_time(dglm_ctx *c, double *t, double *dt, int *tstep, int *done) {
  if (!c) return fail(DGLM_ERR_ARG, "dglm_get_time: null context");
  HIP_TRY(hipSetDevice(c->dev));
  DevScalars s;
  HIP_TRY(hipMemcpyAsync(&s, c->scal, sizeof(s), hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  if (t) *t = s.t_next;
  if (dt) *dt = s.dt;
  if (tstep) *tstep = s.tstep;
  if (done) *done = !(s.t_next < c->p.tmax);
  if (s.bad) return fail(DGLM_ERR_STATE, "non-finite or non-positive dt encountered (state diverged)");
  return DGLM_OK;
}

extern "C" int dglm_rhs(dglm_ctx *c, const double *u, double *rhs, double *maxv, double *tau) {
  if (!c || !u || !rhs) return fail(DGLM_ERR_ARG, "dglm_rhs: null argument");
  if (int rc = need_dirichlet(c, "dglm_rhs")) return rc;
  ws_fn ws;
  st_fn st0, st1;
  if (!dispatch(c, ws, st0, st1)) return fail(DGLM_ERR_UNSUPPORTED, "dglm_rhs: order not built");
  HIP_TRY(hipSetDevice(c->dev));
  const size_t n = (size_t)c->nvar * c->K * c->Np;
  if (!c->rhs) {
    int rc = dev_alloc(&c->rhs, n, "rhs");
    if (rc) return rc;
  }
  HIP_TRY(hipMemcpyAsync(c->u[c->cur], u, n * sizeof(double), hipMemcpyHostToDevice, c->stream));
  ws(c, c->u[c->cur], c->maxv, c->partial, c->partial_idx);
  StageArgs a = stage_args(c, -1);
  st1(c, a);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipMemcpyAsync(rhs, c->rhs, n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  std::vector<double> lam;
  if (maxv || tau) {
    lam.resize((size_t)c->K * 4);
    HIP_TRY(hipMemcpyAsync(lam.data(), c->maxv, lam.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  }
  HIP_TRY(hipStreamSynchronize(c->stream));
  if (maxv) memcpy(maxv, lam.data(), lam.size() * sizeof(double));
  if (tau)
    for (int o = 0; o < c->K * 4; o++) {
      int cc = c->h_conn[o];
      tau[o] = 0.25 * (lam[o] + lam[conn_nbr(cc) * 4 + conn_face(cc)]) * c->p.tau_scale;
    }
  return DGLM_OK;
}

namespace {
// One LSRK4(5) step: wave speed, device-side dt, five fused stages.  The
// wave-speed block maxima go through partial / partial_idx [nblk_ws]; dt_hist
// (device, optional) receives dt at hist_idx.  timed: per-kernel HIP events
// when the context's timing is on.
void enqueue_step(dglm_ctx *c, ws_fn ws, st_fn st0, double *dt_hist, int hist_idx, double *partial,
                  int *partial_idx, bool timed = true) {
  const double cfl_fac = c->p.cfl / (2 * c->p.N + 1);
  int tk = timed ? timer_start(c, 0) : -1;
  ws(c, c->u[c->cur], c->maxv, partial, partial_idx);
  timer_stop(c, tk);
  tk = timed ? timer_start(c, 1) : -1;
  hipLaunchKernelGGL(finalize_dt_kernel, dim3(1), dim3(1024), 0, c->stream, partial, partial_idx, c->nblk_ws,
                     c->scal, cfl_fac, c->p.tmax, dt_hist, (int *)nullptr, (int *)nullptr, hist_idx);
  timer_stop(c, tk);
  for (int s = 0; s < 5; s++) {
    StageArgs a = stage_args(c, s);
    tk = timed ? timer_start(c, 2) : -1;
    st0(c, a);
    timer_stop(c, tk);
    c->cur ^= 1;
  }
}

void drop_graphs(dglm_ctx *c) {
  for (int g = 0; g < 2; g++)
    if (c->graph[g]) {
      hipGraphExecDestroy(c->graph[g]);
      c->graph[g] = nullptr;
    }
}

// Capture two steps (10 stages: the state returns to the same buffer) into
// an executable graph.  Only kernel launches are recorded.
int build_graph(dglm_ctx *c, ws_fn ws, st_fn st0) {
  const int start = c->cur;
  hipGraph_t g = nullptr;
  HIP_TRY(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
  enqueue_step(c, ws, st0, nullptr, 0, c->partial, c->partial_idx);
  enqueue_step(c, ws, st0, nullptr, 0, c->partial, c->partial_idx);
  hipError_t e = hipStreamEndCapture(c->stream, &g);
  c->cur = start;
  if (e != hipSuccess) return fail(DGLM_ERR_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
  e = hipGraphInstantiate(&c->graph[start], g, nullptr, nullptr, 0);
  hipGraphDestroy(g);
  if (e != hipSuccess) {
    c->graph[start] = nullptr;
    return fail(DGLM_ERR_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
  }
  return DGLM_OK;
}
}  // namespace

extern "C" int dglm_step(dglm_ctx *c, int nsteps, double *dt_hist) {
  if (!c || nsteps < 0) return fail(DGLM_ERR_ARG, "dglm_step: bad arguments");
  if (int rc = need_dirichlet(c, "dglm_step")) return rc;
  ws_fn ws;
  st_fn st0, st1;
  if (!dispatch(c, ws, st0, st1)) return fail(DGLM_ERR_UNSUPPORTED, "dglm_step: order not built");
  HIP_TRY(hipSetDevice(c->dev));
  if (dt_hist && nsteps > c->dt_hist_cap) {
    if (c->dt_hist) HIP_TRY(hipFree(c->dt_hist));
    c->dt_hist = nullptr;
    int rc = dev_alloc(&c->dt_hist, (size_t)nsteps, "dt_hist");
    if (rc) return rc;
    c->dt_hist_cap = nsteps;
  }
  int st = 0;
  // Graph replay when nothing per-step has to come back (no dt history, no
  // per-kernel timing).  The first step runs eagerly so launch shapes are
  // fixed before capture.
  const bool graphs = c->use_graph && !dt_hist && !c->timing;
  if (graphs && nsteps >= 3) {
    if (!c->graph[c->cur ^ 1] && !c->graph[c->cur]) {
      enqueue_step(c, ws, st0, nullptr, 0, c->partial, c->partial_idx);
      st++;
    }
    while (nsteps - st >= 2) {
      if (!c->graph[c->cur]) {
        int rc = build_graph(c, ws, st0);
        if (rc) return rc;
      }
      HIP_TRY(hipGraphLaunch(c->graph[c->cur], c->stream));
      st += 2;
    }
  }
  for (; st < nsteps; st++) {
    enqueue_step(c, ws, st0, dt_hist ? c->dt_hist : nullptr, st, c->partial, c->partial_idx);
    if (c->ev_class.size() > 8192) {
      int rc = timer_flush(c);
      if (rc) return rc;
    }
  }
  HIP_TRY(hipGetLastError());
  if (dt_hist) {
    HIP_TRY(hipMemcpyAsync(dt_hist, c->dt_hist, (size_t)nsteps * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
  }
  return DGLM_OK;
}

extern "C" int dglm_step_noalloc(dglm_ctx *c, int nsteps, double *dt_hist) {
  if (!c || nsteps < 0) return fail(DGLM_ERR_ARG, "dglm_step_noalloc: bad arguments");
  if (nsteps > DGLM_NOALLOC_MAX_STEPS)
    return fail(DGLM_ERR_ARG, "dglm_step_noalloc: nsteps = " + std::to_string(nsteps) +
                                  " exceeds DGLM_NOALLOC_MAX_STEPS = " + std::to_string(DGLM_NOALLOC_MAX_STEPS) +
                                  " (step in chunks)");
  if (c->nblk_ws > DGLM_NOALLOC_MAX_BLOCKS)
    return fail(DGLM_ERR_ARG, "dglm_step_noalloc: the mesh has " + std::to_string(c->nblk_ws) +
                                  " wave-speed blocks, more than DGLM_NOALLOC_MAX_BLOCKS = " +
                                  std::to_string(DGLM_NOALLOC_MAX_BLOCKS));
  if (int rc = need_dirichlet(c, "dglm_step_noalloc")) return rc;
  ws_fn ws;
  st_fn st0, st1;
  if (!dispatch(c, ws, st0, st1)) return fail(DGLM_ERR_UNSUPPORTED, "dglm_step_noalloc: order not built");
  HIP_TRY(hipSetDevice(c->dev));
  // Every scratch entry a step reads is written earlier in the same step
  // (block maxima by the wave-speed kernel, dt by the dt kernel), so nothing
  // left by an earlier call is read and nothing needs clearing.
  double *hist = c->scratch;
  double *part = c->scratch + DGLM_NOALLOC_MAX_STEPS;
  int *pidx = (int *)(part + DGLM_NOALLOC_MAX_BLOCKS);
  for (int st = 0; st < nsteps; st++) enqueue_step(c, ws, st0, dt_hist ? hist : nullptr, st, part, pidx, false);
  HIP_TRY(hipGetLastError());
  if (dt_hist) {
    HIP_TRY(hipMemcpyAsync(dt_hist, hist, (size_t)nsteps * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
  }
  return DGLM_OK;
}

extern "C" int dglm_set_graphs(dglm_ctx *c, int on) {
  if (!c) return fail(DGLM_ERR_ARG, "dglm_set_graphs: null context");
  c->use_graph = on != 0;
  if (!c->use_graph) drop_graphs(c);
  return DGLM_OK;
}

extern "C" int dglm_set_stage_kernel(dglm_ctx *c, int variant) {
  if (!c) return fail(DGLM_ERR_ARG, "dglm_set_stage_kernel: null context");
  if (variant != DGLM_STAGE_AUTO && variant != DGLM_STAGE_SIMPLE && variant != DGLM_STAGE_MFMA &&
      variant != DGLM_STAGE_MFMA_WS && variant != DGLM_STAGE_MFMA_MX)
    return fail(DGLM_ERR_ARG, "dglm_set_stage_kernel: unknown variant");
  if (variant == DGLM_STAGE_MFMA_MX && !c->mx_tab)
    return fail(DGLM_ERR_UNSUPPORTED, "dglm_set_stage_kernel: the fused mixed-precision kernel is Euler with "
                                      "precision DGLM_FP32, N <= 6 and no Dirichlet faces");
  if (variant == DGLM_STAGE_MFMA_WS && !c->afrag_ws)
    return fail(DGLM_ERR_UNSUPPORTED,
                "dglm_set_stage_kernel: the warp-specialised kernel is Euler with N = 3 or 4 and precision "
                "DGLM_FP64 or DGLM_FP32");
  if (variant == DGLM_STAGE_MFMA && !mfma_available(c->p.N))
    return fail(DGLM_ERR_UNSUPPORTED, "dglm_set_stage_kernel: no MFMA kernel for this order");
  if (variant == DGLM_STAGE_SIMPLE && c->p.precision != DGLM_FP64)
    return fail(DGLM_ERR_UNSUPPORTED, "dglm_set_stage_kernel: the LDS-staged kernel is FP64 only");
  c->variant = variant;
  drop_graphs(c);  // captured launches name the previous kernel
  return DGLM_OK;
}

extern "C" int dglm_get_stage_kernel(dglm_ctx *c, int *variant) {
  if (!c || !variant) return fail(DGLM_ERR_ARG, "dglm_get_stage_kernel: null argument");
  *variant = stage_kind(c);
  return DGLM_OK;
}

extern "C" int dglm_set_timing(dglm_ctx *c, int on) {
  if (!c) return fail(DGLM_ERR_ARG, "dglm_set_timing: null context");
  int rc = timer_flush(c);
  if (rc) return rc;
  c->timing = on != 0;
  for (int i = 0; i < DGLM_NTIMERS; i++) {
    c->t_ms[i] = 0;
    c->t_cnt[i] = 0;
  }
  return DGLM_OK;
}

extern "C" int dglm_get_timing(dglm_ctx *c, double *ms, long *count, int n) {
  if (!c || !ms || !count) return fail(DGLM_ERR_ARG, "dglm_get_timing: null argument");
  int rc = timer_flush(c);
  if (rc) return rc;
  for (int i = 0; i < n && i < DGLM_NTIMERS; i++) {
    ms[i] = c->t_ms[i];
    count[i] = c->t_cnt[i];
  }
  return DGLM_OK;
}

extern "C" int dglm_integrals(dglm_ctx *c, const double *J, const double *M, double *out) {
  if (!c || !J || !M || !out) return fail(DGLM_ERR_ARG, "dglm_integrals: null argument");
  const int Np = c->Np, K = c->K;
  HIP_TRY(hipSetDevice(c->dev));
  std::vector<double> w(Np, 0.0);
  for (int i = 0; i < Np; i++)
    for (int j = 0; j < Np; j++) w[i] += M[(size_t)j * Np + i];
  const int nblk = (K + 255) / 256;
  int rc;
  if (!c->d_J && ((rc = dev_alloc(&c->d_J, (size_t)K, "J")) || (rc = dev_alloc(&c->d_w, (size_t)Np, "weights")) ||
                  (rc = dev_alloc(&c->d_ipart, (size_t)nblk * c->nvar, "integral partials"))))
    return rc;
  HIP_TRY(hipMemcpyAsync(c->d_J, J, (size_t)K * sizeof(double), hipMemcpyHostToDevice, c->stream));
  HIP_TRY(hipMemcpyAsync(c->d_w, w.data(), (size_t)Np * sizeof(double), hipMemcpyHostToDevice, c->stream));
  hipLaunchKernelGGL(integrals_kernel, dim3(nblk), dim3(256), 0, c->stream, K, Np, c->nvar, c->u[c->cur], c->d_J,
                     c->d_w, c->d_ipart);
  HIP_TRY(hipGetLastError());
  std::vector<double> part((size_t)nblk * c->nvar);
  HIP_TRY(hipMemcpyAsync(part.data(), c->d_ipart, part.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  for (int v = 0; v < c->nvar; v++) {
    double s = 0.0;
    for (int b = 0; b < nblk; b++) s += part[(size_t)b * c->nvar + v];
    out[v] = s;
  }
  return DGLM_OK;
}

extern "C" int dglm_sync(dglm_ctx *c) {
  if (!c) return fail(DGLM_ERR_ARG, "dglm_sync: null context");
  HIP_TRY(hipSetDevice(c->dev));
  HIP_TRY(hipStreamSynchronize(c->stream));
  return DGLM_OK;
}

// ---------------------------------------------------------------------------
// Discrete adjoint: RHS adjoint, trajectory recording with checkpoints, and
// the reverse sweep (see kernels/dg_adjoint.hpp for the per-stage algebra).
//
// Tape layout (all in HBM):
//   ckpt   states at the start of steps 0, E, 2E, ...       (ceil(n/E) states)
//   seg    the 5E+1 stage states of one segment of E steps, recomputed from
//          its checkpoint during the reverse sweep
//   lam    wave speeds per face of each step of the segment (tau)
//   dt, fstar, clip   per step: dt, argmax face of lambda*Fscale, clip flag
// With E ~ sqrt(n) this holds ~sqrt(n) + 5 sqrt(n) states, and the forward is
// recomputed once.
// ---------------------------------------------------------------------------
namespace {

typedef void (*adj_fn)(dglm_ctx *, const AdjArgs &);
typedef void (*lam_fn)(dglm_ctx *, const double *, const double *, int);

template <int N, int PHYS>
void launch_adj_a(dglm_ctx *c, const AdjArgs &a) {
  constexpr int EB = eb_adjoint<N, PHYS>();
  hipLaunchKernelGGL((adjoint_pass_a<N, PHYS>), dim3((c->K + EB - 1) / EB), dim3(256), 0, c->stream, a);
}

template <int N>
void launch_adj_a_mfma(dglm_ctx *c, const AdjArgs &a);

template <int N, int PHYS>
void launch_adj_b(dglm_ctx *c, const AdjArgs &a) {
  const int n = c->K * Ord<N>::Np;
  hipLaunchKernelGGL((adjoint_pass_b<N, PHYS>), dim3((n + 255) / 256), dim3(256), 0, c->stream, a);
}

template <int N>
void launch_lam(dglm_ctx *c, const double *u, const double *lam, int step);

template <int N, int PHYS>
int adj_blocks() {
  constexpr int EB = eb_adjoint<N, PHYS>();
  return EB;
}

// The MFMA pass A (FP64 Euler, N <= 4) where it fits a workgroup.
template <int PHYS>
bool pick_adj_mfma(int N, adj_fn &A, int &eb) {
  if (PHYS != PHYS_EULER) return false;
  switch (N) {
#define DGLM_MCASE(n)                          \
  case n:                                      \
    if (!adj_mfma_available<n>()) return false; \
    A = launch_adj_a_mfma<n>;                  \
    eb = AdjShape<n>::EB;                      \
    return true;
    DGLM_MCASE(1)
    DGLM_MCASE(2)
    DGLM_MCASE(3)
    DGLM_MCASE(4)
#undef DGLM_MCASE
    default:
      return false;
  }
}

template <int PHYS>
bool pick_adj_impl(int N, adj_fn &A, adj_fn &B, lam_fn &Lf, int &eb);

template <int PHYS>
bool pick_adj(int N, adj_fn &A, adj_fn &B, lam_fn &Lf, int &eb) {
  if (!pick_adj_impl<PHYS>(N, A, B, Lf, eb)) return false;
  pick_adj_mfma<PHYS>(N, A, eb);
  return true;
}

template <int PHYS>
bool pick_adj_impl(int N, adj_fn &A, adj_fn &B, lam_fn &Lf, int &eb) {
  switch (N) {
#define DGLM_ACASE(n)                      \
  case n:                                  \
    A = launch_adj_a<n, PHYS>;             \
    B = launch_adj_b<n, PHYS>;             \
    Lf = launch_lam<n>;                    \
    eb = adj_blocks<n, PHYS>();            \
    return true;
    DGLM_ACASE(1)
    DGLM_ACASE(2)
    DGLM_ACASE(3)
    DGLM_ACASE(4)
    DGLM_ACASE(5)
    DGLM_ACASE(6)
    DGLM_ACASE(7)
    DGLM_ACASE(8)
    DGLM_ACASE(9)
    DGLM_ACASE(10)
#undef DGLM_ACASE
    default:
      return false;
  }
}

StageArgs make_stage_args(dglm_ctx *c, const double *u_in, double *u_out, double *k, const double *maxv,
                          const DevScalars *scal, int s) {
  StageArgs a = stage_args(c, s);
  a.u_in = u_in;
  a.u_out = u_out;
  a.k = k;
  a.maxv = maxv;
  a.scal = scal;
  return a;
}

}  // namespace

struct dglm_adj_tape {
  int nsteps = 0, every = 0, nck = 0, eb = 1, nblk_a = 0;
  double *afragT = nullptr;  // A^T fragments of the MFMA pass A (null: scalar pass A)
  double *ckpt = nullptr, *ck_t = nullptr, *dt = nullptr, *seg = nullptr, *lam = nullptr, *kscr = nullptr;
  int *fstar = nullptr, *clip = nullptr, *ck_step = nullptr;
  DevScalars *scal2 = nullptr;
  double *ubar = nullptr, *kbar = nullptr, *gsur = nullptr, *taubar = nullptr, *dtpart = nullptr, *tsc = nullptr;
  // second face-term buffer of the MFMA pass A, which takes over the pass B
  // of stages 4..1 (each pass A reads the terms of the one before it while
  // writing its own)
  double *gsur2 = nullptr;
  double *rbar = nullptr;
  int *slots = nullptr, *fpos = nullptr;
  // The last nkeep segments kept whole: their stage states [5*every + 1][nst]
  // and wave speeds [every][K*4] per segment, written by the recording
  // forward itself, so the reverse sweep does not recompute them.
  double *keep = nullptr, *keep_lam = nullptr;
  int nkeep = 0;
  size_t bytes = 0, keep_bytes = 0;
};

namespace {

template <int N>
void launch_adj_a_mfma(dglm_ctx *c, const AdjArgs &a) {
  if constexpr (adj_mfma_available<N>()) {
    using S = AdjShape<N>;
    hipLaunchKernelGGL((adjoint_pass_a_mfma<N>), dim3((c->K + S::EB - 1) / S::EB), dim3(S::NT), 0, c->stream, a,
                       c->tape->afragT);
  }
}

template <int N>
void launch_lam(dglm_ctx *c, const double *u, const double *lam, int step) {
  dglm_adj_tape *T = c->tape;
  constexpr int EB = lam_adj_elems(Ord<N>::Np);
  hipLaunchKernelGGL((lambda_adjoint_kernel<N>), dim3((c->K + EB - 1) / EB), dim3(4 * EB), 0, c->stream, c->K, u, lam,
                     T->taubar, c->conn, c->fmx, T->dt, T->fstar, step, T->tsc, c->p.tau_scale, c->p.gamma,
                     T->ubar);
}

void free_tape(dglm_adj_tape *T) {
  if (!T) return;
  void *ptrs[] = {T->keep, T->keep_lam, T->ckpt, T->ck_t, T->dt, T->seg, T->lam, T->kscr, T->fstar, T->clip, T->ck_step, T->scal2,
                  T->ubar, T->kbar, T->gsur, T->gsur2, T->taubar, T->dtpart, T->tsc, T->rbar, T->slots, T->fpos,
                  T->afragT};
  for (void *q : ptrs)
    if (q) hipFree(q);
  delete T;
}

// Buffers common to the RHS adjoint and the reverse sweep.
int ensure_adjoint_buffers(dglm_ctx *c) {
  if (!c->tape) c->tape = new dglm_adj_tape();
  dglm_adj_tape *T = c->tape;
  if (T->ubar) return DGLM_OK;
  adj_fn A, B;
  lam_fn L;
  int eb = 1;
  bool ok = c->p.physics == DGLM_EULER ? pick_adj<PHYS_EULER>(c->p.N, A, B, L, eb)
                                       : pick_adj<PHYS_ADV>(c->p.N, A, B, L, eb);
  if (!ok) return fail(DGLM_ERR_UNSUPPORTED, "adjoint: order not built");
  T->eb = eb;
  T->nblk_a = (c->K + eb - 1) / eb;
  const size_t nst = (size_t)c->nvar * c->K * c->Np;
  // face-node slots of each volume node, and node positions in each face mask
  RefOps ref;
  reference_operators(c->p.N, ref);
  const int Np = c->Np, Nfp = c->Nfp;
  std::vector<int> slots((size_t)Np * 3, -1), fpos((size_t)4 * Np, -1);
  for (int m = 0; m < 4 * Nfp; m++) {
    const int node = ref.Fmask[m];
    for (int q = 0; q < 3; q++)
      if (slots[node * 3 + q] < 0) {
        slots[node * 3 + q] = m;
        break;
      }
    fpos[(m / Nfp) * Np + node] = m % Nfp;
  }
  int rc;
  if ((rc = dev_alloc(&T->ubar, nst, "ubar")) || (rc = dev_alloc(&T->kbar, nst, "kbar")) ||
      (rc = dev_alloc(&T->gsur, (size_t)c->nvar * c->K * c->NFF, "face adjoint")) ||
      (rc = dev_alloc(&T->taubar, (size_t)c->K * 4, "taubar")) ||
      (rc = dev_alloc(&T->dtpart, (size_t)T->nblk_a, "dt partials")) || (rc = dev_alloc(&T->tsc, 4, "t scalars")) ||
      (rc = dev_alloc(&T->slots, slots.size(), "node slots")) || (rc = dev_alloc(&T->fpos, fpos.size(), "fpos")))
    return rc;
  HIP_TRY(hipMemcpy(T->slots, slots.data(), slots.size() * sizeof(int), hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(T->fpos, fpos.data(), fpos.size() * sizeof(int), hipMemcpyHostToDevice));
  T->bytes += (2 * nst + (size_t)c->nvar * c->K * c->NFF + (size_t)c->K * 4 + T->nblk_a) * sizeof(double);
  // A^T fragments for the MFMA pass A
  adj_fn Am;
  int ebm = 0;
  if (c->p.physics == DGLM_EULER && pick_adj_mfma<PHYS_EULER>(c->p.N, Am, ebm)) {
    std::vector<double> opT((size_t)c->KD * Np), frag;
    HIP_TRY(hipMemcpy(opT.data(), c->op, opT.size() * sizeof(double), hipMemcpyDeviceToHost));
    switch (c->p.N) {
#define DGLM_FCASE(n)                                                            \
  case n:                                                                        \
    frag.resize((size_t)AdjShape<n>::KT * AdjShape<n>::KSA * 64);               \
    adj_arrange<n>(opT.data(), frag.data());                                     \
    break;
      DGLM_FCASE(1)
      DGLM_FCASE(2)
      DGLM_FCASE(3)
      DGLM_FCASE(4)
#undef DGLM_FCASE
      default:
        break;
    }
    if ((rc = dev_alloc(&T->afragT, frag.size(), "A^T fragments"))) return rc;
    HIP_TRY(hipMemcpy(T->afragT, frag.data(), frag.size() * sizeof(double), hipMemcpyHostToDevice));
    const size_t ng = (size_t)c->nvar * c->K * c->NFF;
    if ((rc = dev_alloc(&T->gsur2, ng, "face adjoint (second)"))) return rc;
    T->bytes += ng * sizeof(double);
  }
  return DGLM_OK;
}

AdjArgs adj_args(dglm_ctx *c) {
  dglm_adj_tape *T = c->tape;
  AdjArgs a{};
  a.K = c->K;
  a.ubar = T->ubar;
  a.kbar = T->kbar;
  a.gsur = T->gsur;
  a.taubar = T->taubar;
  a.vel = c->vel;
  a.geo = c->geo;
  a.conn = c->conn;
  a.fmx = c->fmx;
  a.node_slots = T->slots;
  a.fpos = T->fpos;
  a.op = c->op;
  a.gamma = c->p.gamma;
  a.tau_scale = c->p.tau_scale;
  return a;
}

}  // namespace

extern "C" int dglm_adjoint_rhs(dglm_ctx *c, const double *u, const double *rbar, double *ubar, double *taubar) {
  if (!c || !u || !rbar || !ubar) return fail(DGLM_ERR_ARG, "dglm_adjoint_rhs: null argument");
  if (c->p.flux != DGLM_FLUX_LM) return fail(DGLM_ERR_UNSUPPORTED, "adjoint: built for the LM flux only");
  if (c->nd > 0) return fail(DGLM_ERR_UNSUPPORTED, "adjoint: not built for meshes with Dirichlet faces");
  HIP_TRY(hipSetDevice(c->dev));
  int rc = ensure_adjoint_buffers(c);
  if (rc) return rc;
  dglm_adj_tape *T = c->tape;
  const size_t nst = (size_t)c->nvar * c->K * c->Np;
  if (!T->rbar && (rc = dev_alloc(&T->rbar, nst, "rbar"))) return rc;
  ws_fn ws;
  st_fn st0, st1;
  adj_fn A, B;
  lam_fn L;
  int eb;
  if (!dispatch(c, ws, st0, st1)) return fail(DGLM_ERR_UNSUPPORTED, "dglm_adjoint_rhs: order not built");
  if (c->p.physics == DGLM_EULER) pick_adj<PHYS_EULER>(c->p.N, A, B, L, eb);
  else pick_adj<PHYS_ADV>(c->p.N, A, B, L, eb);
  HIP_TRY(hipMemcpyAsync(c->u[c->cur], u, nst * sizeof(double), hipMemcpyHostToDevice, c->stream));
  HIP_TRY(hipMemcpyAsync(T->rbar, rbar, nst * sizeof(double), hipMemcpyHostToDevice, c->stream));
  ws(c, c->u[c->cur], c->maxv, c->partial, c->partial_idx);  // tau at u
  AdjArgs a = adj_args(c);
  a.mode = 1;
  a.u = c->u[c->cur];
  a.rbar = T->rbar;
  a.maxv = c->maxv;
  A(c, a);
  B(c, a);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipMemcpyAsync(ubar, T->ubar, nst * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  if (taubar)
    HIP_TRY(hipMemcpyAsync(taubar, T->taubar, (size_t)c->K * 4 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  return DGLM_OK;
}

extern "C" int dglm_adjoint_record(dglm_ctx *c, int nsteps, int every) {
  if (!c || nsteps <= 0 || every <= 0) return fail(DGLM_ERR_ARG, "dglm_adjoint_record: bad arguments");
  if (c->p.flux != DGLM_FLUX_LM) return fail(DGLM_ERR_UNSUPPORTED, "adjoint: built for the LM flux only");
  if (c->nd > 0) return fail(DGLM_ERR_UNSUPPORTED, "adjoint: not built for meshes with Dirichlet faces");
  HIP_TRY(hipSetDevice(c->dev));
  int rc = ensure_adjoint_buffers(c);
  if (rc) return rc;
  dglm_adj_tape *T = c->tape;
  ws_fn ws;
  st_fn st0, st1;
  if (!dispatch(c, ws, st0, st1)) return fail(DGLM_ERR_UNSUPPORTED, "dglm_adjoint_record: order not built");
  const size_t nst = (size_t)c->nvar * c->K * c->Np;
  const int nck = (nsteps + every - 1) / every;
  // (re)allocate the recording buffers when the shape changes
  if (T->nsteps != nsteps || T->every != every) {
    // The tape is empty until every buffer of the new shape exists: a failed
    // allocation leaves nsteps = 0, which dglm_adjoint_backward refuses, and a
    // retry with any shape reallocates.  Kept segments go first (the largest
    // buffers), so the new shape's buffers do not compete with them.
    T->nsteps = T->every = T->nck = 0;
    T->nkeep = -1;
    void *old[] = {T->keep, T->keep_lam, T->ckpt, T->ck_t, T->dt, T->seg, T->lam, T->kscr, T->fstar, T->clip,
                   T->ck_step, T->scal2};
    for (void *q : old)
      if (q) HIP_TRY(hipFree(q));
    T->keep = T->keep_lam = nullptr;
    T->keep_bytes = 0;
    T->ckpt = T->ck_t = T->dt = T->seg = T->lam = T->kscr = nullptr;
    T->fstar = T->clip = T->ck_step = nullptr;
    T->scal2 = nullptr;
    if ((rc = dev_alloc(&T->ckpt, nst * nck, "checkpoints")) || (rc = dev_alloc(&T->ck_t, (size_t)nck, "ckpt t")) ||
        (rc = dev_alloc(&T->ck_step, (size_t)nck, "ckpt step")) || (rc = dev_alloc(&T->dt, (size_t)nsteps, "dt")) ||
        (rc = dev_alloc(&T->fstar, (size_t)nsteps, "fstar")) || (rc = dev_alloc(&T->clip, (size_t)nsteps, "clip")) ||
        (rc = dev_alloc(&T->seg, nst * (5 * (size_t)every + 1), "segment states")) ||
        (rc = dev_alloc(&T->lam, (size_t)c->K * 4 * every, "segment wave speeds")) ||
        (rc = dev_alloc(&T->kscr, nst, "k scratch")) || (rc = dev_alloc(&T->scal2, 1, "recompute scalars")))
      return rc;
    T->nsteps = nsteps;
    T->every = every;
    T->nck = nck;
    T->bytes = (2 * nst + (size_t)c->nvar * c->K * c->NFF * (T->gsur2 ? 2 : 1) + (size_t)c->K * 4 + T->nblk_a) *
                   sizeof(double) +
               (nst * (nck + 5 * (size_t)every + 2) + (size_t)c->K * 4 * every) * sizeof(double);
  }
  // Segments kept whole: HBM bought back from the recompute.  The count is the
  // caller's cap (dglm_adjoint_set_keep) or, by default, what half of the free
  // device memory holds; sized once per tape shape, and again at every
  // recording that holds none (the memory may have become free since).
  const size_t seg_n = nst * (5 * (size_t)every + 1), lam_n = (size_t)c->K * 4 * every;
  if (T->nkeep < 0 || (T->nkeep == 0 && c->adj_keep_max != 0) ||
      (c->adj_keep_max >= 0 && T->nkeep > c->adj_keep_max)) {
    if (T->keep) HIP_TRY(hipFree(T->keep));
    if (T->keep_lam) HIP_TRY(hipFree(T->keep_lam));
    T->keep = T->keep_lam = nullptr;
    T->bytes -= T->keep_bytes;
    T->keep_bytes = 0;
    size_t fr = 0, tot = 0;
    HIP_TRY(hipMemGetInfo(&fr, &tot));
    const size_t per = (seg_n + lam_n) * sizeof(double);
    long nk = (long)((fr / 2) / per);
    if (c->adj_keep_max >= 0 && nk > c->adj_keep_max) nk = c->adj_keep_max;
    if (nk > nck) nk = nck;
    T->nkeep = (int)nk;
    if (T->nkeep > 0) {
      if (device_malloc((void **)&T->keep, seg_n * T->nkeep * sizeof(double) + 16) != hipSuccess ||
          device_malloc((void **)&T->keep_lam, lam_n * T->nkeep * sizeof(double) + 16) != hipSuccess) {
        // the budget is an estimate: without the memory, keep nothing and
        // recompute every segment (clear the allocation error first)
        (void)hipGetLastError();
        if (T->keep) hipFree(T->keep);
        if (T->keep_lam) hipFree(T->keep_lam);
        T->keep = T->keep_lam = nullptr;
        T->nkeep = 0;
      } else {
        T->keep_bytes = per * T->nkeep;
        T->bytes += T->keep_bytes;
      }
    }
  }
  const int q0 = nck - T->nkeep;  // first kept segment
  const double cfl_fac = c->p.cfl / (2 * c->p.N + 1);
  for (int n = 0; n < nsteps; n++) {
    const int q = n / every;
    if (n % every == 0) {
      HIP_TRY(hipMemcpyAsync(T->ckpt + nst * q, c->u[c->cur], nst * sizeof(double), hipMemcpyDeviceToDevice,
                             c->stream));
      HIP_TRY(hipMemcpyAsync(T->ck_t + q, &c->scal->t_next, sizeof(double), hipMemcpyDeviceToDevice, c->stream));
      HIP_TRY(hipMemcpyAsync(T->ck_step + q, &c->scal->tstep, sizeof(int), hipMemcpyDeviceToDevice, c->stream));
    }
    if (q >= q0) {
      // a kept segment: the stages write straight into its tape slots (the
      // same kernels and inputs as the reverse sweep's recompute)
      const int n0 = q * every, n1 = std::min(n0 + every, nsteps);
      double *base = T->keep + seg_n * (q - q0) + nst * 5 * (size_t)(n - n0);
      double *lam = T->keep_lam + lam_n * (q - q0) + (size_t)c->K * 4 * (n - n0);
      if (n == n0)
        HIP_TRY(hipMemcpyAsync(base, c->u[c->cur], nst * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
      ws(c, base, lam, c->partial, c->partial_idx);
      hipLaunchKernelGGL(finalize_dt_kernel, dim3(1), dim3(1024), 0, c->stream, c->partial, c->partial_idx,
                         c->nblk_ws, c->scal, cfl_fac, c->p.tmax, T->dt, T->fstar, T->clip, n);
      for (int s = 0; s < 5; s++) st0(c, make_stage_args(c, base + nst * s, base + nst * (s + 1), c->k, lam, c->scal, s));
      if (n == n1 - 1)  // the context's state continues from the segment's last stage
        HIP_TRY(hipMemcpyAsync(c->u[c->cur], base + nst * 5, nst * sizeof(double), hipMemcpyDeviceToDevice,
                               c->stream));
      continue;
    }
    ws(c, c->u[c->cur], c->maxv, c->partial, c->partial_idx);
    hipLaunchKernelGGL(finalize_dt_kernel, dim3(1), dim3(1024), 0, c->stream, c->partial, c->partial_idx,
                       c->nblk_ws, c->scal, cfl_fac, c->p.tmax, T->dt, T->fstar, T->clip, n);
    for (int s = 0; s < 5; s++) {
      StageArgs a = stage_args(c, s);
      st0(c, a);
      c->cur ^= 1;
    }
  }
  HIP_TRY(hipGetLastError());
  return DGLM_OK;
}

extern "C" int dglm_adjoint_set_keep(dglm_ctx *c, int max_segments) {
  if (!c) return fail(DGLM_ERR_ARG, "dglm_adjoint_set_keep: null context");
  c->adj_keep_max = max_segments < 0 ? -1 : max_segments;
  if (c->tape) c->tape->nkeep = -1;  // resized at the next recording
  return DGLM_OK;
}

extern "C" int dglm_adjoint_backward(dglm_ctx *c, const double *ubar_T, double *ubar_0, int device_ptrs) {
  if (!c || !ubar_T || !ubar_0) return fail(DGLM_ERR_ARG, "dglm_adjoint_backward: null argument");
  dglm_adj_tape *T = c->tape;
  if (!T || T->nsteps == 0) return fail(DGLM_ERR_STATE, "dglm_adjoint_backward: nothing recorded");
  if (c->nd > 0) return fail(DGLM_ERR_UNSUPPORTED, "adjoint: not built for meshes with Dirichlet faces");
  HIP_TRY(hipSetDevice(c->dev));
  ws_fn ws;
  st_fn st0, st1;
  adj_fn A, B;
  lam_fn L;
  int eb;
  if (!dispatch(c, ws, st0, st1)) return fail(DGLM_ERR_UNSUPPORTED, "dglm_adjoint_backward: order not built");
  if (c->p.physics == DGLM_EULER) pick_adj<PHYS_EULER>(c->p.N, A, B, L, eb);
  else pick_adj<PHYS_ADV>(c->p.N, A, B, L, eb);
  const size_t nst = (size_t)c->nvar * c->K * c->Np;
  const bool exact = c->adj_exact && c->p.physics == DGLM_EULER;
  const double cfl_fac = c->p.cfl / (2 * c->p.N + 1);
  const hipMemcpyKind kin = device_ptrs ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
  HIP_TRY(hipMemcpyAsync(T->ubar, ubar_T, nst * sizeof(double), kin, c->stream));
  HIP_TRY(hipMemsetAsync(T->kbar, 0, nst * sizeof(double), c->stream));
  HIP_TRY(hipMemsetAsync(T->kscr, 0, nst * sizeof(double), c->stream));
  HIP_TRY(hipMemsetAsync(T->tsc, 0, 4 * sizeof(double), c->stream));
  const int q0 = T->nck - std::max(T->nkeep, 0);  // first segment kept whole by the recording
  for (int q = T->nck - 1; q >= 0; q--) {
    const int n0 = q * T->every, n1 = std::min(n0 + T->every, T->nsteps);
    double *seg = T->seg, *segl = T->lam;
    if (q >= q0) {
      seg = T->keep + nst * (5 * (size_t)T->every + 1) * (q - q0);
      segl = T->keep_lam + (size_t)c->K * 4 * T->every * (q - q0);
    } else {
      // recompute the segment's stage states from its checkpoint
      HIP_TRY(hipMemcpyAsync(T->seg, T->ckpt + nst * q, nst * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
      HIP_TRY(hipMemsetAsync(T->scal2, 0, sizeof(DevScalars), c->stream));
      HIP_TRY(hipMemcpyAsync(&T->scal2->t_next, T->ck_t + q, sizeof(double), hipMemcpyDeviceToDevice, c->stream));
      for (int n = n0; n < n1; n++) {
        double *base = T->seg + nst * 5 * (size_t)(n - n0);
        double *lam = T->lam + (size_t)c->K * 4 * (n - n0);
        ws(c, base, lam, c->partial, c->partial_idx);
        hipLaunchKernelGGL(finalize_dt_kernel, dim3(1), dim3(1024), 0, c->stream, c->partial, c->partial_idx,
                           c->nblk_ws, T->scal2, cfl_fac, c->p.tmax, (double *)nullptr, (int *)nullptr,
                           (int *)nullptr, 0);
        for (int s = 0; s < 5; s++)
          st0(c, make_stage_args(c, base + nst * s, base + nst * (s + 1), T->kscr, lam, T->scal2, s));
      }
    }
    // reverse sweep over the segment
    for (int n = n1 - 1; n >= n0; n--) {
      double *base = seg + nst * 5 * (size_t)(n - n0);
      const double *lam = segl + (size_t)c->K * 4 * (n - n0);
      HIP_TRY(hipMemsetAsync(T->taubar, 0, (size_t)c->K * 4 * sizeof(double), c->stream));
      HIP_TRY(hipMemsetAsync(T->dtpart, 0, (size_t)T->nblk_a * sizeof(double), c->stream));
      for (int s = 4; s >= 0; s--) {
        AdjArgs a = adj_args(c);
        a.mode = 0;
        if (T->gsur2) {
          // the MFMA pass A adds the partner terms of the stage before it
          // (written to the other buffer); only stage 0's go through pass B,
          // ahead of the wave-speed adjoint
          a.gsur = (s & 1) ? T->gsur2 : T->gsur;
          a.gsur_in = s < 4 ? ((s & 1) ? T->gsur : T->gsur2) : nullptr;
        }
        a.u = base + nst * s;
        a.maxv = lam;
        a.dtp = T->dt + n;
        a.a = RKA[s];
        a.b = RKB[s];
        if (exact) {
          a.u_next = base + nst * (s + 1);
          a.u_prev = s > 0 ? base + nst * (s - 1) : nullptr;
          a.a_prev = s > 0 ? RKA[s - 1] : 0.0;
          a.b_prev = s > 0 ? RKB[s - 1] : 1.0;
          a.dtpart = T->dtpart;
        }
        A(c, a);
        if (!T->gsur2 || s == 0) B(c, a);
      }
      if (exact) {
        hipLaunchKernelGGL(adjoint_dt_reduce, dim3(1), dim3(256), 0, c->stream, T->dtpart, T->nblk_a, T->clip, n,
                           T->tsc);
        L(c, base, lam, n);
      }
    }
  }
  HIP_TRY(hipGetLastError());
  const hipMemcpyKind kout = device_ptrs ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost;
  HIP_TRY(hipMemcpyAsync(ubar_0, T->ubar, nst * sizeof(double), kout, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  // The kept segments have served their sweep: return their memory (the
  // largest part of the tape) to the device.  A repeated sweep of the same
  // recording recomputes every segment from its checkpoint; the next
  // recording sizes the kept segments anew.
  if (T->keep || T->keep_lam) {
    if (T->keep) HIP_TRY(hipFree(T->keep));
    if (T->keep_lam) HIP_TRY(hipFree(T->keep_lam));
    T->keep = T->keep_lam = nullptr;
    T->bytes -= T->keep_bytes;
    T->keep_bytes = 0;
    T->nkeep = -1;
  }
  return DGLM_OK;
}

extern "C" int dglm_adjoint_set_exact(dglm_ctx *c, int on) {
  if (!c) return fail(DGLM_ERR_ARG, "dglm_adjoint_set_exact: null context");
  c->adj_exact = on != 0;
  return DGLM_OK;
}

extern "C" int dglm_adjoint_tape_bytes(dglm_ctx *c, double *bytes) {
  if (!c || !bytes) return fail(DGLM_ERR_ARG, "dglm_adjoint_tape_bytes: null argument");
  *bytes = c->tape ? (double)c->tape->bytes : 0.0;
  return DGLM_OK;
}
