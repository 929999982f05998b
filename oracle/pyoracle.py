"""ctypes wrapper of liboracle.so and runner of the compiled reference (_ref/).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, never by the product package.
"""
import ctypes as C
import math
import os
import subprocess
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
REF_DIR = os.path.join(HERE, "_ref")

PD = C.POINTER(C.c_double)
PI = C.POINTER(C.c_int)


class OracleProblem(C.Structure):
    _fields_ = ([("N", C.c_int), ("Np", C.c_int), ("Nfp", C.c_int), ("K", C.c_int), ("nvar", C.c_int),
                 ("cfl", C.c_double), ("gamma", C.c_double), ("tau_scale", C.c_double), ("tmax", C.c_double)]
                + [(n, PI) for n in ("vmapP", "vmapF", "Fmask")]
                + [(n, PD) for n in ("WDr", "WDs", "WDt", "LIFT", "Fscale", "rx", "ry", "rz", "sx", "sy", "sz",
                                     "tx", "ty", "tz", "nx", "ny", "nz", "vx", "vy", "vz")]
                + [("EtoB", PI), ("gD", PD), ("flux", C.c_int)])


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            raise ImportError(f"{LIB} missing: run `make -C oracle oracle`")
        _lib = C.CDLL(LIB)
    return _lib


def _d(a):
    return a.ctypes.data_as(PD)


def _i(a):
    return a.ctypes.data_as(PI)


class Oracle:
    """CPU restatement bound to one problem (datain-style arrays)."""

    def __init__(self, problem, physics, cfl=0.8, gamma=1.4, tau_scale=1.2, tmax=math.inf, velocity=None,
                 flux="lm"):
        """flux: 'lm' (the reference's) or 'rusanov' (Rusanov for Euler, exact
        upwind for advection)."""
        a = problem.arrays
        self._keep = dict(a)
        self.nvar = 5 if physics == "euler" else 1
        self.K, self.Np = problem.K, problem.Np
        P = OracleProblem()
        P.N, P.Np, P.Nfp, P.K, P.nvar = problem.N, problem.Np, problem.Nfp, problem.K, self.nvar
        P.cfl, P.gamma, P.tau_scale, P.tmax = cfl, gamma, tau_scale, tmax
        P.flux = {"lm": 0, "rusanov": 1}[flux]
        for n in ("vmapP", "vmapF", "Fmask"):
            setattr(P, n, _i(a[n]))
        for n in ("WDr", "WDs", "WDt", "LIFT", "Fscale", "rx", "ry", "rz", "sx", "sy", "sz", "tx", "ty", "tz", "nx",
                  "ny", "nz"):
            setattr(P, n, _d(a[n]))
        if self.nvar == 1:
            n = problem.K * problem.Np
            if velocity is None:
                velocity = (1.0, 1.0, 1.0)
            v = [np.ascontiguousarray(np.broadcast_to(np.asarray(c, dtype=np.float64), (n,))) for c in velocity]
            self._keep["vel"] = v
            P.vx, P.vy, P.vz = _d(v[0]), _d(v[1]), _d(v[2])
        if "EtoB" in a:
            self._keep["EtoB"] = np.ascontiguousarray(a["EtoB"], dtype=np.int32)
            P.EtoB = _i(self._keep["EtoB"])
        self.P = P

    def set_dirichlet(self, g):
        """Face states ub on the faces with EtoB == 1, [nvar][K*4*Nfp]
        (the other entries are not read)."""
        g = np.ascontiguousarray(g, dtype=np.float64).reshape(self.nvar, -1)
        self._keep["gD"] = g
        self.P.gD = _d(g)

    def rhs(self, u):
        u = np.ascontiguousarray(u, dtype=np.float64).reshape(self.nvar, -1)
        out = np.empty_like(u)
        maxv = np.empty((self.K, 4))
        tau = np.empty((self.K, 4))
        if lib().oracle_rhs(C.byref(self.P), _d(u), _d(out), _d(maxv), _d(tau)) != 0:
            raise MemoryError("oracle_rhs allocation failed")
        return out, maxv, tau

    def run(self, u, nsteps, t=0.0, k=None):
        u = np.array(u, dtype=np.float64).reshape(self.nvar, -1)
        k = np.zeros_like(u) if k is None else np.array(k, dtype=np.float64)
        tt = C.c_double(t)
        ts = C.c_int(0)
        dts = np.zeros(nsteps)
        if lib().oracle_run(C.byref(self.P), _d(u), _d(k), nsteps, C.byref(tt), C.byref(ts), _d(dts)) != 0:
            raise MemoryError("oracle_run allocation failed")
        return u, dts, tt.value


# ---------------------------------------------------------------------------
# Compiled reference solver (oracle/_ref), this container only.
# ---------------------------------------------------------------------------
def ref_available(physics, N):
    return os.path.exists(os.path.join(REF_DIR, f"{'euler' if physics == 'euler' else 'advec'}_N{N}"))


def write_datain(problem, d, matrix_order="c"):
    os.makedirs(d, exist_ok=True)
    np.array([problem.K], np.int32).tofile(os.path.join(d, "dim.bin"))
    for name, arr in problem.arrays.items():
        arr = np.asarray(arr)
        if name in ("WDr", "WDs", "WDt", "LIFT"):
            arr.ravel(matrix_order).tofile(os.path.join(d, f"{name}.bin"))
        else:
            arr.ravel().tofile(os.path.join(d, f"{name}.bin"))


class RefSolver:
    """Runs oracle/_ref/{advec,euler}_N<n> on a problem written to a temp dir."""

    def __init__(self, problem, physics, cfl=0.8, gamma=1.4, tau_scale=1.2, tmax=1000.0, velocity=None,
                 threads=8):
        self.exe = os.path.join(REF_DIR, f"{'euler' if physics == 'euler' else 'advec'}_N{problem.N}")
        self.nvar = 5 if physics == "euler" else 1
        self.K, self.Np = problem.K, problem.Np
        self.tmp = tempfile.TemporaryDirectory(prefix="dglm_ref_")
        self.datain = os.path.join(self.tmp.name, "datain") + "/"
        self.dataout = os.path.join(self.tmp.name, "dataout") + "/"
        write_datain(problem, self.datain)
        os.makedirs(self.dataout, exist_ok=True)
        self.env = dict(os.environ, DGLM_DATAIN=self.datain, DGLM_DATAOUT=self.dataout, DGLM_TMAX=repr(tmax),
                        DGLM_CFL=repr(cfl), DGLM_GAMMA=repr(gamma), DGLM_TAU_SCALE=repr(tau_scale),
                        OMP_NUM_THREADS=str(threads))
        if velocity is not None and self.nvar == 1:
            n = problem.K * problem.Np
            v = np.concatenate([np.broadcast_to(np.asarray(c, dtype=np.float64), (n,)) for c in velocity])
            vf = os.path.join(self.tmp.name, "vel.bin")
            v.tofile(vf)
            self.env["DGLM_VEL"] = vf

    def _call(self, *args):
        r = subprocess.run([self.exe, *map(str, args)], env=self.env, stdout=subprocess.PIPE, stderr=subprocess.PIPE,
                           text=True)
        if r.returncode != 0:
            raise RuntimeError(f"reference run failed: {r.stderr}")
        return r.stdout

    def rhs(self, u):
        f = os.path.join(self.tmp.name, "u.bin")
        np.ascontiguousarray(u, dtype=np.float64).tofile(f)
        o = os.path.join(self.tmp.name, "rhs.bin")
        self._call("rhs", f, o)
        d = np.fromfile(o)
        n = self.nvar * self.K * self.Np
        return d[:n].reshape(self.nvar, -1), d[n:n + 4 * self.K].reshape(self.K, 4), d[n + 4 * self.K:].reshape(self.K, 4)

    def steps(self, u0, nsteps):
        o = os.path.join(self.tmp.name, "out.bin")
        if u0 is None:
            self._call("steps", nsteps, "-", o)
        else:
            f = os.path.join(self.tmp.name, "u0.bin")
            np.ascontiguousarray(u0, dtype=np.float64).tofile(f)
            self._call("steps", nsteps, f, o)
        d = np.fromfile(o)
        n = self.nvar * self.K * self.Np
        u = d[:n].reshape(self.nvar, -1)
        t, cnt = d[n], int(d[n + 1])
        return u, d[n + 2:n + 2 + cnt], t

    def run(self):
        """The reference main loop to TMAX; returns stdout and the dataout dir."""
        return self._call("run"), self.dataout


def rhs_tl(oracle, u, du, tau=None):
    """Tangent-linear RHS of the oracle with tau frozen (at the given [K, 4]
    penalty, else at the one u gives)."""
    u = np.ascontiguousarray(u, dtype=np.float64).reshape(oracle.nvar, -1)
    du = np.ascontiguousarray(du, dtype=np.float64).reshape(oracle.nvar, -1)
    out = np.empty_like(u)
    tp = None
    if tau is not None:
        tau = np.ascontiguousarray(tau, dtype=np.float64)
        tp = _d(tau)
    rc = lib().oracle_rhs_tl(C.byref(oracle.P), _d(u), _d(du), tp, _d(out))
    if rc == -2:
        raise ValueError("the oracle's tangent-linear RHS is of the LM flux only")
    if rc != 0:
        raise MemoryError("oracle_rhs_tl allocation failed")
    return out


def rhs_tau(oracle, u, tau):
    """Oracle RHS at u with a given face penalty tau [K, 4]."""
    u = np.ascontiguousarray(u, dtype=np.float64).reshape(oracle.nvar, -1)
    tau = np.ascontiguousarray(tau, dtype=np.float64)
    out = np.empty_like(u)
    if lib().oracle_rhs_tau(C.byref(oracle.P), _d(u), _d(tau), _d(out)) != 0:
        raise MemoryError("oracle_rhs_tau allocation failed")
    return out
