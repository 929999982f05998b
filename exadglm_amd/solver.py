"""Python handle on one solver context of libexadglm_hip.so."""
import ctypes as C
import math

import numpy as np

from . import _capi as A

_PHYS = {"advection": A.DGLM_ADVECTION, "euler": A.DGLM_EULER}
_PREC = {"fp64": A.DGLM_FP64, "fp32": A.DGLM_FP32, "bf16x3": A.DGLM_BF16X3, "bf16x6": A.DGLM_BF16X6}


def make_mesh_struct(arrays, K, matrix_order="c"):
    """dglm_mesh view of a datain-style dict (the arrays must outlive the call)."""
    m = A.Mesh()
    m.K = K
    m.matrix_order = 0 if matrix_order == "c" else 1
    for n in ("EtoB", "vmapP", "vmapF", "Fmask"):
        setattr(m, n, A.iptr(arrays[n]) if n in arrays else None)
    for n in ("WDr", "WDs", "WDt", "LIFT", "Fscale", "rx", "ry", "rz", "sx", "sy", "sz", "tx", "ty", "tz", "nx",
              "ny", "nz", "px", "py", "pz"):
        setattr(m, n, A.dptr(arrays[n]) if n in arrays else None)
    return m


class Solver:
    """Owns a device context: state, LSRK register and the step loop.

    physics: 'advection' or 'euler'; precision: 'fp64' (default), or a mixed
    mode for N <= 6 ('fp32', 'bf16x3', 'bf16x6': the contraction of the
    flux remainders after a least-squares affine fit, in FP32 or split bf16).
    tmax: the run ends exactly at tmax (use math.inf for fixed step counts).
    """

    def __init__(self, problem=None, physics="euler", N=None, cfl=0.8, gamma=1.4, tau_scale=1.2, tmax=math.inf,
                 precision="fp64", device=0, arrays=None, K=None, datain_dir=None, matrix_order="c", flux="lm"):
        L = A.lib()
        p = A.Params()
        p.N = N if N is not None else problem.N
        p.physics = _PHYS[physics]
        p.precision = _PREC[precision]
        p.flux = {"lm": A.DGLM_FLUX_LM, "rusanov": A.DGLM_FLUX_RUSANOV}[flux]
        p.cfl, p.gamma, p.tau_scale, p.tmax = cfl, gamma, tau_scale, tmax
        p.device = device
        self.params = p
        self.physics = physics
        self._ctx = C.c_void_p()
        if datain_dir is not None:
            A.check(L.dglm_create_from_dir(C.byref(p), datain_dir.encode(), 0 if matrix_order == "c" else 1,
                                           C.byref(self._ctx)))
        else:
            arrays = arrays if arrays is not None else problem.arrays
            K = K if K is not None else problem.K
            m = make_mesh_struct(arrays, K, matrix_order)
            A.check(L.dglm_create(C.byref(p), C.byref(m), C.byref(self._ctx)))
        K_, Np_, Nfp_, nv_ = C.c_int(), C.c_int(), C.c_int(), C.c_int()
        A.check(L.dglm_sizes(self._ctx, C.byref(K_), C.byref(Np_), C.byref(Nfp_), C.byref(nv_)))
        self.K, self.Np, self.Nfp, self.nvar = K_.value, Np_.value, Nfp_.value, nv_.value
        self.recording = 0  # bumped by every adjoint_record (the tape holds the latest one)

    @property
    def shape(self):
        return (self.nvar, self.K * self.Np)

    def close(self):
        if self._ctx:
            A.lib().dglm_destroy(self._ctx)
            self._ctx = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _state_arg(self, u):
        u = np.ascontiguousarray(u, dtype=np.float64)
        if u.size != self.nvar * self.K * self.Np:
            raise ValueError(f"state must have {self.nvar}x{self.K * self.Np} entries, got {u.shape}")
        return u

    def set_state(self, u):
        u = self._state_arg(u)
        A.check(A.lib().dglm_set_state(self._ctx, A.dptr(u)))

    def get_state(self):
        u = np.empty(self.shape)
        A.check(A.lib().dglm_get_state(self._ctx, A.dptr(u)))
        return u

    def set_state_device(self, ptr):
        A.check(A.lib().dglm_set_state_device(self._ctx, C.cast(C.c_void_p(ptr), A.PD)))

    def get_state_device(self, ptr):
        A.check(A.lib().dglm_get_state_device(self._ctx, C.cast(C.c_void_p(ptr), A.PD)))

    def set_velocity(self, vx, vy, vz):
        n = self.K * self.Np
        v = [np.ascontiguousarray(np.broadcast_to(np.asarray(a, dtype=np.float64), (n,))) for a in (vx, vy, vz)]
        A.check(A.lib().dglm_set_velocity(self._ctx, A.dptr(v[0]), A.dptr(v[1]), A.dptr(v[2])))

    def dirichlet_faces(self):
        """Face indices e*4 + f of the Dirichlet faces (EtoB == 1) in
        ascending order: the slot order of set_dirichlet."""
        nd = C.c_int()
        A.check(A.lib().dglm_dirichlet_faces(self._ctx, C.byref(nd), None))
        faces = np.empty(nd.value, np.int32)
        if nd.value:
            A.check(A.lib().dglm_dirichlet_faces(self._ctx, C.byref(nd), A.iptr(faces)))
        return faces

    def set_dirichlet(self, g):
        """Prescribed boundary states on the Dirichlet faces: g [nvar][nd][Nfp]
        (slots as dirichlet_faces()), or a face-node array [nvar][K*4*Nfp]
        (vmapM order) whose Dirichlet entries are taken.  LM flux: the face
        state ub; Rusanov: the exterior trace u+."""
        g = np.asarray(g, dtype=np.float64)
        faces = self.dirichlet_faces()
        nd, nfp = len(faces), self.Nfp
        if g.size == self.nvar * self.K * 4 * nfp and g.size != self.nvar * nd * nfp:
            g = g.reshape(self.nvar, self.K * 4, nfp)[:, faces, :]
        if g.size != self.nvar * nd * nfp:
            raise ValueError(f"Dirichlet states must have {self.nvar}x{nd}x{nfp} entries, got {g.shape}")
        g = np.ascontiguousarray(g.reshape(self.nvar, nd, nfp))
        A.check(A.lib().dglm_set_dirichlet(self._ctx, A.dptr(g)))

    def set_dirichlet_device(self, ptr):
        A.check(A.lib().dglm_set_dirichlet_device(self._ctx, C.cast(C.c_void_p(ptr), A.PD)))

    def set_time(self, t=0.0, tstep=0):
        A.check(A.lib().dglm_set_time(self._ctx, C.c_double(t), tstep))

    def time(self):
        t, dt, ts, done = C.c_double(), C.c_double(), C.c_int(), C.c_int()
        A.check(A.lib().dglm_get_time(self._ctx, C.byref(t), C.byref(dt), C.byref(ts), C.byref(done)))
        return t.value, dt.value, ts.value, bool(done.value)

    def rhs(self, u):
        u = self._state_arg(u)
        out = np.empty(self.shape)
        maxv = np.empty((self.K, 4))
        tau = np.empty((self.K, 4))
        A.check(A.lib().dglm_rhs(self._ctx, A.dptr(u), A.dptr(out), A.dptr(maxv), A.dptr(tau)))
        return out, maxv, tau

    def step(self, nsteps=1, return_dt=False):
        if return_dt:
            dts = np.empty(nsteps)
            A.check(A.lib().dglm_step(self._ctx, nsteps, A.dptr(dts)))
            return dts
        A.check(A.lib().dglm_step(self._ctx, nsteps, None))
        return None

    def step_noalloc(self, nsteps=1, return_dt=False):
        """step() through the allocation-free entry point (no device
        allocation by the library; scratch in its fixed per-device array)."""
        dts = np.empty(nsteps) if return_dt else None
        A.check(A.lib().dglm_step_noalloc(self._ctx, nsteps, A.dptr(dts)))
        return dts

    def set_stage_kernel(self, variant):
        """'auto', 'simple' (LDS-staged FP64), 'mfma' (one workgroup per batch,
        operator rows in registers) or 'mfma_ws' (persistent warp-specialised
        FP64 Euler kernel, N = 3, 4)."""
        code = {"auto": A.DGLM_STAGE_AUTO, "simple": A.DGLM_STAGE_SIMPLE, "mfma": A.DGLM_STAGE_MFMA,
                "mfma_ws": A.DGLM_STAGE_MFMA_WS, "mfma_mx": A.DGLM_STAGE_MFMA_MX}[variant]
        A.check(A.lib().dglm_set_stage_kernel(self._ctx, code))

    def stage_kernel(self):
        v = C.c_int()
        A.check(A.lib().dglm_get_stage_kernel(self._ctx, C.byref(v)))
        return {A.DGLM_STAGE_SIMPLE: "simple", A.DGLM_STAGE_MFMA: "mfma", A.DGLM_STAGE_MFMA_WS: "mfma_ws",
                A.DGLM_STAGE_MFMA_MX: "mfma_mx"}[v.value]

    def set_timing(self, on=True):
        A.check(A.lib().dglm_set_timing(self._ctx, 1 if on else 0))

    def timing(self):
        """{'wavespeed'|'dt'|'stage': (total_ms, launches)} since set_timing(True)."""
        ms = (C.c_double * 3)()
        cnt = (C.c_long * 3)()
        A.check(A.lib().dglm_get_timing(self._ctx, ms, cnt, 3))
        return {k: (ms[i], cnt[i]) for i, k in enumerate(("wavespeed", "dt", "stage"))}

    def set_graphs(self, on=True):
        A.check(A.lib().dglm_set_graphs(self._ctx, 1 if on else 0))

    def sync(self):
        A.check(A.lib().dglm_sync(self._ctx))

    def integrals(self, J, M):
        out = np.empty(self.nvar)
        J = np.ascontiguousarray(J, dtype=np.float64)
        M = np.ascontiguousarray(M, dtype=np.float64)
        A.check(A.lib().dglm_integrals(self._ctx, A.dptr(J), A.dptr(M), A.dptr(out)))
        return out

    # ---- discrete adjoint ----
    def adjoint_rhs(self, u, rbar):
        """(dR/du)^T rbar at u (tau held at u) and the per-face tau adjoint."""
        u = self._state_arg(u)
        rbar = self._state_arg(rbar)
        ubar = np.empty(self.shape)
        taubar = np.empty((self.K, 4))
        A.check(A.lib().dglm_adjoint_rhs(self._ctx, A.dptr(u), A.dptr(rbar), A.dptr(ubar), A.dptr(taubar)))
        return ubar, taubar

    def adjoint_record(self, nsteps, every=None):
        """Advance nsteps from the current state, checkpointing every `every`
        steps (default ~sqrt(nsteps))."""
        if every is None:
            every = max(1, int(round(math.sqrt(nsteps))))
        self.recording += 1  # a failed recording also invalidates the previous tape
        A.check(A.lib().dglm_adjoint_record(self._ctx, nsteps, every))

    def adjoint_backward(self, ubar_T):
        ubar_T = self._state_arg(ubar_T)
        out = np.empty(self.shape)
        A.check(A.lib().dglm_adjoint_backward(self._ctx, A.dptr(ubar_T), A.dptr(out), 0))
        return out

    def adjoint_backward_device(self, ubar_T_ptr, ubar_0_ptr):
        A.check(A.lib().dglm_adjoint_backward(self._ctx, C.cast(C.c_void_p(ubar_T_ptr), A.PD),
                                              C.cast(C.c_void_p(ubar_0_ptr), A.PD), 1))

    def set_adjoint_exact(self, on=True):
        A.check(A.lib().dglm_adjoint_set_exact(self._ctx, 1 if on else 0))

    def set_adjoint_keep(self, max_segments=-1):
        """Checkpoint segments the recording keeps whole (no recompute in the
        reverse sweep): -1 = as many as half the free device memory holds, 0 = none.
        Their memory is held from the recording to the end of the next reverse
        sweep."""
        A.check(A.lib().dglm_adjoint_set_keep(self._ctx, int(max_segments)))

    def tape_bytes(self):
        b = C.c_double()
        A.check(A.lib().dglm_adjoint_tape_bytes(self._ctx, C.byref(b)))
        return b.value
