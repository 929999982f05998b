/*
 * exadglm.h -- C ABI of the MI355X explicit DGLM solver core (libexadglm_hip.so).
 *
 * Plain C types only: host pointers, sizes and status codes.  Every function
 * returns DGLM_OK (0) on success or a negative DGLM_ERR_* code; the message of
 * the last failure on the calling thread is returned by dglm_last_error().
 * No function aborts the host process.
 *
 * Array conventions (identical to the reference's datain/ file contract,
 * dim3/src/python/dglm_3d_launcher.py:74-149):
 *   volume-node arrays     [K*Np]          element-major, u[v][e*Np + i] for NVAR fields
 *   face arrays            [K*4]           index e*4 + f
 *   face-node arrays       [K*4*Nfp]       index (e*4 + f)*Nfp + j
 *   operator matrices      row-major [Np*Np] / [Np*4*Nfp] unless matrix_order says 'f'
 */
#ifndef EXADGLM_H
#define EXADGLM_H

#ifdef __cplusplus
extern "C" {
#endif

enum {
  DGLM_OK = 0,
  DGLM_ERR_ARG = -1,         /* invalid argument or shape */
  DGLM_ERR_HIP = -2,         /* HIP runtime failure (message names the call) */
  DGLM_ERR_IO = -3,          /* missing / truncated datain file */
  DGLM_ERR_STATE = -4,       /* non-finite state or dt detected on the device */
  DGLM_ERR_UNSUPPORTED = -5  /* order / physics / connectivity not built in */
};

enum { DGLM_ADVECTION = 0, DGLM_EULER = 1 };

/* Arithmetic of the operator contractions.  The LSRK state, pressure from
 * energy, wave speed and conservation sums are FP64 in every mode.  The
 * mixed modes remove a least-squares affine fit of every flux block in FP64
 * and contract only the remainder in low precision: DGLM_FP32 on FP32 MFMA;
 * DGLM_BF16X3 on bf16 MFMA as the split-BF16 three-term product
 * (hi*hi + hi*lo + lo*hi of two bf16 pieces per operand); DGLM_BF16X6 with
 * three pieces and six products.  FP32 accumulation in all three.  Each
 * mode promises a result within its bound of FP64 (DGLM_FP32 and
 * DGLM_BF16X6 1e-8 relative), not a particular kernel: at Euler N = 3, 4
 * DGLM_FP32 runs the FP64 warp-specialised kernel, which is faster there than
 * the mixed-precision one (dglm_set_stage_kernel(DGLM_STAGE_MFMA_MX) still
 * selects the latter). */
enum { DGLM_FP64 = 0, DGLM_FP32 = 1, DGLM_BF16X3 = 2, DGLM_BF16X6 = 3 };

/* Numerical flux.  DGLM_FLUX_LM is the reference's Lagrange-multiplier flux
 * (n.F(ub) - tau (ub - u-)) Fscale, dim3/src/cpp/euler_dglm.cpp:110-183.
 * DGLM_FLUX_RUSANOV is the textbook local Lax-Friedrichs flux
 * (n.F(u-) + n.F(u+))/2 + lambda/2 (u- - u+), lambda = max(|v.n| + c) over the
 * two traces (exact upwinding for advection); it has no reference
 * counterpart and no adjoint. */
enum { DGLM_FLUX_LM = 0, DGLM_FLUX_RUSANOV = 1 };

const char *dglm_last_error(void);
int dglm_version(void);

/* Testing aid (no reference counterpart): the next n device allocations made
 * by the library fail as if the device were out of memory, so that callers'
 * tests can exercise its allocation-failure paths without exhausting HBM.
 * Process-wide; 0 clears it. */
int dglm_set_alloc_faults(int n);

/* ---------------------------------------------------------------------------
 * Reference-element tables.  Replaces TetraRef / Operator3DRef / Operator3DFace
 * (dim3/src/python/tetra.py:305-331, operator_3d_ref.py:302-349,
 * operator_3d_face.py:275-301), which read mpmath-generated tables from
 * dim3/data/bin_mats; here they are computed in extended precision.
 * ------------------------------------------------------------------------- */
int dglm_ref_sizes(int N, int *Np, int *Nfp);

typedef struct dglm_ref_tables {
  double *xyz, *rst;         /* [Np*3] equilateral and reference coordinates */
  double *V, *M, *invM;      /* [Np*Np] Vandermonde, mass, inverse mass       */
  double *Dr, *Ds, *Dt;      /* [Np*Np] strong differentiation                */
  double *WDr, *WDs, *WDt;   /* [Np*Np] weak differentiation M^-1 D^T M       */
  double *Emat, *LIFT;       /* [Np*4*Nfp] face mass, M^-1 Emat               */
  int *Fmask;                /* [4*Nfp] face node indices (vmapM / LIFT order) */
  int *FmaskP;               /* [4*3*Nfp] neighbour order per rotation        */
} dglm_ref_tables;

/* Any NULL member is skipped. */
int dglm_ref_operators(int N, const dglm_ref_tables *out);

/* Interpolation to order high_N nodes and the order-high_N mass matrix used by
 * the L2 error (operator_3d_ref.py:265-299, dglm_3d_launcher.py:328-355).
 * high_rst [hNp*3], interpM [hNp*Np], high_M [hNp*hNp]. */
int dglm_interp_operators(int N, int high_N, double *high_rst, double *interpM, double *high_M);

/* GLL nodes and orthonormal Jacobi values (jacobi.py:53-114). */
int dglm_gll_nodes(int N, double *x /*[N+1]*/);
int dglm_jacobi(int n, double a, double b, int npts, const double *x, double *P, double *dP);

/* ---------------------------------------------------------------------------
 * Mesh.  Replaces Uniform3DSingle (uniform_3d_single.py:8-134 with the cores
 * update_EtoV / update_EtoEFRB, uniform_3d_single_core.cpp:8-277) and the
 * serial parts of Mesh3D (mesh_3d.py:9-116; mesh_3d_core.cpp:10-356, 468-505).
 * ------------------------------------------------------------------------- */
/* box = {x1,x2,y1,y2,z1,z2}; phy_tags[a] < 0 means periodic along axis a,
 * otherwise the physical tag stored in EtoB on that axis' outer faces.
 * VXYZ [nx*ny*nz*3]; EtoV, EtoE, EtoF, EtoR, EtoB [K*4], K = 6(nx-1)(ny-1)(nz-1). */
int dglm_uniform_mesh(int nx, int ny, int nz, const double *box, const int *phy_tags, double *VXYZ, int *EtoV,
                      int *EtoE, int *EtoF, int *EtoR, int *EtoB);
int dglm_mesh_points(int K, int Np, const int *EtoV, const double *VXYZ, const double *rst, double *PX, double *PY,
                     double *PZ);
/* metrics [9*K] = rx,ry,rz,sx,sy,sz,tx,ty,tz (each [K]) */
int dglm_mesh_metrics(int K, const int *EtoV, const double *VXYZ, double *J, double *metrics);
int dglm_mesh_normals(int K, const double *J, const double *metrics, double *nx, double *ny, double *nz, double *sJ,
                      double *Fscale);
int dglm_mesh_vmaps(int K, int N, const int *EtoE, const int *EtoF, const int *EtoR, int *vmapM, int *vmapP,
                    int *vmapF);
int dglm_mesh_spatial_h(int K, const int *EtoV, const double *VXYZ, double *h);

/* ---------------------------------------------------------------------------
 * Solver.  Replaces the compile-time composed executable of
 * dim3/src/cpp/{advec,euler}_{data,dglm,time,init}.cpp and
 * dim3/src/cuda/{advec,euler}_nopad_*.cu: DataHost/DataDev construction
 * (euler_data.cpp:71-178, euler_nopad_data_dev.cu:46-204) and runRK
 * (euler_time.cpp:41-107, euler_nopad_time.cu:32-108).  The #define knobs of
 * the generated common.h (dglm_3d_launcher.py:152-205) become dglm_params.
 * ------------------------------------------------------------------------- */
typedef struct dglm_params {
  int N;           /* polynomial order (NP, NFP follow): 1..10 in FP64,  */
                   /* 1..6 in mixed precision                           */
  int physics;     /* DGLM_ADVECTION (NVAR=1) or DGLM_EULER (NVAR=5)     */
  int precision;   /* DGLM_FP64 / DGLM_FP32 / DGLM_BF16X3 / DGLM_BF16X6 */
  int flux;        /* DGLM_FLUX_LM or DGLM_FLUX_RUSANOV                  */
  double cfl;      /* CFL                                                */
  double gamma;    /* GAMMA                                              */
  double tau_scale;/* TAU_SCALE                                          */
  double tmax;     /* TMAX; the last step is clipped to end exactly here */
  int device;      /* HIP device ordinal (DEVID)                         */
} dglm_params;

/* The datain/ array set (names as the files; shapes as the header above). */
typedef struct dglm_mesh {
  int K;
  int matrix_order;  /* 0: WD*, LIFT row-major ('c'); 1: column-major ('f') */
  const int *EtoB, *vmapP, *vmapF, *Fmask;
  const double *WDr, *WDs, *WDt, *LIFT;
  const double *Fscale;
  const double *rx, *ry, *rz, *sx, *sy, *sz, *tx, *ty, *tz;
  const double *nx, *ny, *nz;
  const double *px, *py, *pz;
} dglm_mesh;

typedef struct dglm_ctx dglm_ctx;

int dglm_create(const dglm_params *p, const dglm_mesh *m, dglm_ctx **out);
/* Reads WORK_DIR/datain/{dim,EtoB,vmapP,vmapF,Fmask,WDr,...,pz}.bin as written
 * by save_to_binaries with dtype 'f8' or 'f4' (REAL = double or float,
 * dglm_3d_launcher.py:138-141,154; the width follows from the file sizes and
 * f4 values are widened exactly); a missing file, a file of the wrong size or
 * f4 files mixed with f8 ones is DGLM_ERR_IO (the reference calls
 * std::terminate, utils/src/cpp/io.cpp:7-17). */
int dglm_create_from_dir(const dglm_params *p, const char *datain_dir, int matrix_order, dglm_ctx **out);
int dglm_destroy(dglm_ctx *ctx);

int dglm_sizes(const dglm_ctx *ctx, int *K, int *Np, int *Nfp, int *nvar);
/* u: host [nvar][K*Np].  Setting the state zeroes the LSRK register k. */
int dglm_set_state(dglm_ctx *ctx, const double *u);
int dglm_get_state(dglm_ctx *ctx, double *u);
/* Device-pointer variants (same layout) for callers that own HBM buffers. */
int dglm_set_state_device(dglm_ctx *ctx, const double *d_u);
int dglm_get_state_device(dglm_ctx *ctx, double *d_u);
/* Advection velocity per node (vx, vy, vz each [K*Np]); advec_init.cpp:5-23. */
int dglm_set_velocity(dglm_ctx *ctx, const double *vx, const double *vy, const double *vz);

/* Dirichlet faces: the faces with EtoB == 1, which must be boundary faces
 * (vmapP = vmapM there).  The reference's calc_fluxLM has the branch as a
 * stub, "ub = function of x,y,z,t" (euler_dglm.cpp:136-141,
 * advec_dglm.cpp:113-119), and leaves ub unset; here the caller prescribes
 * the boundary state g per face node.  With the LM flux it is the face state
 * ub (the tau term is tau (g - u-)); with the Rusanov option it is the
 * exterior trace u+.
 *
 * dglm_dirichlet_faces: nd = their number; faces (optional, [nd]) = their
 * face indices e*4 + f in ascending order, the slot order of g.
 * dglm_set_dirichlet: g host [nvar][nd][Nfp], node j of a face in the order
 * of its vmapM entries (Fmask[f]).  The states are held until set again:
 * constant over the stages and steps of a dglm_step call (step one at a time
 * and set g between calls for time-dependent data).  dglm_step / dglm_rhs on
 * a mesh with Dirichlet faces return DGLM_ERR_ARG until g has been set once.
 * The adjoint refuses such meshes (DGLM_ERR_UNSUPPORTED). */
int dglm_dirichlet_faces(const dglm_ctx *ctx, int *nd, int *faces);
int dglm_set_dirichlet(dglm_ctx *ctx, const double *g);
int dglm_set_dirichlet_device(dglm_ctx *ctx, const double *d_g);

int dglm_set_time(dglm_ctx *ctx, double t, int tstep);
/* done != 0 once t has reached tmax. */
int dglm_get_time(dglm_ctx *ctx, double *t, double *dt, int *tstep, int *done);

/* One RHS evaluation at state u (host [nvar][K*Np]) with the wave speed and
 * tau taken from u itself: calc_maxv, update_tau, update_ub, calc_fluxLM,
 * calc_fluxes, update_rhs (euler_dglm.cpp:6-302).  maxv and tau ([K*4]) are
 * optional outputs. */
int dglm_rhs(dglm_ctx *ctx, const double *u, double *rhs, double *maxv, double *tau);

/* Advance nsteps LSRK4(5) steps (dt adaptive per step, computed on the device).
 * dt_hist [nsteps] may be NULL.  Steps after t reaches tmax are no-ops. */
int dglm_step(dglm_ctx *ctx, int nsteps, double *dt_hist);

/* Allocation-free stepping.  No counterpart in the reference: it is for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  Advances nsteps like
 * dglm_step, with the same kernels and bitwise the same states and dt history
 * (dt_hist [nsteps], host, may be NULL), but calls no hipMalloc, hipFree or
 * allocating library routine: its device scratch -- the dt history and the
 * wave-speed block maxima -- lives in a fixed array of the library
 * (DGLM_NOALLOC_SCRATCH_BYTES on every device it runs on, held until the
 * process ends), and the steps are launched one kernel at a time (no graph,
 * no timing events).  nsteps > DGLM_NOALLOC_MAX_STEPS, or a mesh with more
 * than DGLM_NOALLOC_MAX_BLOCKS wave-speed blocks (K/64 at N <= 4, K/32 at
 * N = 5, 6, K/16 above), is DGLM_ERR_ARG.  All calls on one device share that
 * array: calls that overlap on one device (several threads or contexts at
 * once) race. */
#define DGLM_NOALLOC_MAX_STEPS 65536
#define DGLM_NOALLOC_MAX_BLOCKS 262144
#define DGLM_NOALLOC_SCRATCH_BYTES (8 * (DGLM_NOALLOC_MAX_STEPS + DGLM_NOALLOC_MAX_BLOCKS + DGLM_NOALLOC_MAX_BLOCKS / 2))
int dglm_step_noalloc(dglm_ctx *ctx, int nsteps, double *dt_hist);

/* Conserved integrals sum_e J_e 1^T M u_e per variable, FP64 on the device
 * with a fixed-order reduction; J [K], M [Np*Np]; out [nvar]. */
int dglm_integrals(dglm_ctx *ctx, const double *J, const double *M, double *out);

/* dglm_step replays a captured two-step hipGraph when no dt history or
 * timing is requested (default on). */
int dglm_set_graphs(dglm_ctx *ctx, int on);

int dglm_sync(dglm_ctx *ctx);

/* Per-kernel timing with HIP events on the context's stream (the reference
 * profiles with Nsight Compute, utils/src/python/gpu_profile.py:98-310).
 * Turning timing on resets the totals.  dglm_get_timing returns, for kernel
 * classes 0 = wave speed, 1 = dt finalize, 2 = fused stage, the summed
 * milliseconds and launch counts (synchronises the stream). */
int dglm_set_timing(dglm_ctx *ctx, int on);

/* Stage-kernel selection (the reference picks its GPU variant by file prefix,
 * *_nopad_ / *_pad_, dglm_3d_launcher.py:176-205).  SIMPLE: LDS-staged FP64;
 * MFMA: one workgroup per 16-element batch, operator row tiles in registers
 * (N <= 6, FP64 and mixed precision); MFMA_WS: persistent warp-specialised
 * FP64 Euler kernel (producer waves stage the next batch while consumer waves
 * run the MFMA contraction; N = 3, 4, precision DGLM_FP64 or DGLM_FP32);
 * MFMA_MX: fused mixed-precision Euler kernel (DGLM_FP32, N <= 6, no
 * Dirichlet faces).  AUTO takes MFMA_WS where it exists, else MFMA_MX, else
 * MFMA, else SIMPLE (N = 7..10).  dglm_get_stage_kernel reports the kernel
 * that runs. */
enum { DGLM_STAGE_AUTO = 0, DGLM_STAGE_SIMPLE = 1, DGLM_STAGE_MFMA = 2, DGLM_STAGE_MFMA_WS = 3, DGLM_STAGE_MFMA_MX = 4 };
int dglm_set_stage_kernel(dglm_ctx *ctx, int variant);
int dglm_get_stage_kernel(dglm_ctx *ctx, int *variant);
int dglm_get_timing(dglm_ctx *ctx, double *ms, long *count, int n);

/* ---------------------------------------------------------------------------
 * Discrete adjoint (no counterpart in the reference, which has no
 * differentiation; it differentiates the path above exactly as implemented:
 * update_rhs, calc_fluxLM, calc_fluxes and lserk_stage of euler_dglm.cpp /
 * euler_time.cpp, including tau(u) and dt(u) through the argmax the forward
 * took).  All adjoint arithmetic is FP64; a mixed-precision forward is
 * recomputed with the same kernels during the reverse sweep.
 * ------------------------------------------------------------------------- */
/* ubar = (dR/du)^T rbar at state u with tau held at u's wave speeds; taubar
 * ([K*4], optional) = dR/dtau^T rbar per face side.  Host arrays. */
int dglm_adjoint_rhs(dglm_ctx *ctx, const double *u, const double *rbar, double *ubar, double *taubar);
/* Advance nsteps from the current state like dglm_step, keeping a checkpoint
 * every `every` steps (sqrt(nsteps) balances memory and recompute). */
int dglm_adjoint_record(dglm_ctx *ctx, int nsteps, int every);
/* Reverse sweep over the recorded steps: ubar_0 = (du_T/du_0)^T ubar_T.
 * device_ptrs != 0: both pointers are device memory ([nvar][K*Np]).
 * DGLM_ERR_STATE when nothing is recorded (also after a failed recording). */
int dglm_adjoint_backward(dglm_ctx *ctx, const double *ubar_T, double *ubar_0, int device_ptrs);
/* on = 0 freezes tau and dt at their forward values (the exact gradient of
 * the frozen-coefficient map); default 1 (exact; only Euler depends on it). */
int dglm_adjoint_set_exact(dglm_ctx *ctx, int on);
/* Checkpoint segments kept whole on the tape: the recording forward writes
 * the stage states and wave speeds of the last `max_segments` segments
 * straight into device memory, and the reverse sweep skips their recompute
 * (the gradient is bitwise the same).  -1 (default): as many as half of the
 * free device memory holds at the recording; 0: none.  Takes effect at the
 * next dglm_adjoint_record.  That memory is held from the recording to the end
 * of the next dglm_adjoint_backward, which releases it (a second sweep of the
 * same recording recomputes every segment).  If it cannot be allocated the
 * recording keeps no segment (the next recording tries again).  No reference counterpart (the reference has no
 * adjoint). */
int dglm_adjoint_set_keep(dglm_ctx *ctx, int max_segments);
/* Device memory of the adjoint buffers and tape, kept segments included. */
int dglm_adjoint_tape_bytes(dglm_ctx *ctx, double *bytes);

#ifdef __cplusplus
}
#endif

#endif /* EXADGLM_H */
