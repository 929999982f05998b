"""Drop-in driver with the reference launcher's API (dim3/src/python/dglm_3d_launcher.py).

The reference composes a C++/CUDA executable per problem (save_to_header,
link_codes, build) and runs it as a child process that reads WORK_DIR/datain
and writes WORK_DIR/dataout plus "tstep=.., dt=.., t=.." lines (:260-325).
Here the same calls drive the HIP library in-process:

  dg = DGLM3DSingleGPU(dglm_dir, work_dir, N, Nvar, tmax, dglm_param, devid=0)
  dg.init_mesh({"type": "uniform", "param": [nx, ny, nz, x1, x2, y1, y2, z1, z2, bc]})
  dg.make_work_dir(); dg.save_to_binaries()
  dg.save_to_header([...]); dg.link_codes([...]); dg.build([...])   # recorded, no compile
  tstep, avg_dt, elapsed = dg.run("euler_nopad_main")
  l2 = dg.l2_error(u_ref, u_numeric)

run() reads the datain/ files just written (the same file contract), sets the
reference initial condition (advec_init.cpp / euler_init.cpp) unless an
initial state was given with set_initial_state(), and writes dataout/ at the
reference cadence (find_print_tstep, euler_time.cpp:4-12).  Both the CPU and
the GPU classes run on the GPU: the library has no CPU solver.
"""
import json
import os
import re
import time

import numpy as np

from . import _capi as A
from .mesh import Mesh3D, Uniform3DSingle
from .problem import ic_advection_sine, ic_euler_density_wave
from .reference import Operator3DFace, Operator3DRef, TetraRef
from .solver import Solver


def find_print_tstep(tmax, dt, base_num_tstep):
    """Output cadence: leading digit of int(tmax/dt)/base times its power of ten."""
    interval = int(tmax / dt) // base_num_tstep
    s = str(interval)
    return int(s[0]) * 10 ** (len(s) - 1)


class DGLM3DBase:
    def __init__(self, dglm_dir, work_dir, N, Nvar, tmax, dglm_param):
        self.dglm_dir = os.path.abspath(dglm_dir) if dglm_dir else None
        self.work_dir = os.path.abspath(work_dir)
        self.N = N
        self.Nvar = Nvar
        self.tmax = tmax
        self.param = {"CFL": 0.8, "gamma": 1.4, "tau_scale": 1.2, "interpN": 2, "dtype": "f8"}
        self.param.update(dglm_param)
        if self.param["dtype"] not in ("f8", "f4"):
            raise ValueError("dtype must be 'f8' or 'f4'")
        self.tet = TetraRef(N, self.param["interpN"])
        self.op_ref = Operator3DRef(self.tet)
        self.op_face = Operator3DFace(self.op_ref)
        self.Np, self.Nfp, self.Nface = self.tet.Np, self.tet.Nfp, self.tet.Nface
        self.header = {}
        self.filenames = []
        self._u0 = None
        self._velocity = None
        self._dirichlet = None
        self.devid = 0

    def init_mesh(self, mesh_info=None, verbose=1):
        pass

    def make_work_dir(self):
        self.datain_dir = os.path.join(self.work_dir, "datain")
        self.dataout_dir = os.path.join(self.work_dir, "dataout")
        os.makedirs(self.datain_dir, exist_ok=True)
        os.makedirs(self.dataout_dir, exist_ok=True)

    def save_to_binaries(self, extra_arr_dict=None, matrix_order="c", verbose=1):
        """Write datain/*.bin in the reference's layout (launcher :74-149)."""
        m = self.mesh
        arr = {"dim": np.array([m.Nelem], "i4"), "EtoB": m.EtoB, "vmapP": m.vmapP, "vmapF": m.vmapF,
               "Fmask": self.op_face.Fmask, "WDr": self.op_ref.WDr, "WDs": self.op_ref.WDs,
               "WDt": self.op_ref.WDt, "LIFT": self.op_face.LIFT, "Fscale": m.Fscale, "rx": m.rx, "ry": m.ry,
               "rz": m.rz, "sx": m.sx, "sy": m.sy, "sz": m.sz, "tx": m.tx, "ty": m.ty, "tz": m.tz, "nx": m.NX,
               "ny": m.NY, "nz": m.NZ, "px": m.PX, "py": m.PY, "pz": m.PZ}
        arr.update(extra_arr_dict or {})
        if verbose >= 1:
            print(f"Save DG arrays to *.bin files in {self.datain_dir.replace(self.work_dir, 'WORK_DIR')}/")
        for name, a in arr.items():
            a = np.asarray(a)
            if a.dtype in (np.float32, np.float64):
                a = a.astype(self.param["dtype"])
            path = os.path.join(self.datain_dir, f"{name}.bin")
            with open(path, "wb") as f:
                (a.ravel(matrix_order) if name in ("WDr", "WDs", "WDt", "LIFT") else a.ravel()).tofile(f)
        self.matrix_order = matrix_order

    def save_to_header(self, filename_list=None, extra_str_dict=None, extra_num_dict=None, verbose=1):
        """Record the problem constants the reference bakes into common.h (:152-205)."""
        self.header = {"REAL": {"f8": "double", "f4": "float"}[self.param["dtype"]], "N": self.N, "NP": self.Np,
                       "NFP": self.Nfp, "NFACE": 4, "NROT": 3, "NVAR": self.Nvar, "TMAX": self.tmax,
                       "CFL": self.param["CFL"], "GAMMA": self.param["gamma"], "TAU_SCALE": self.param["tau_scale"]}
        self.header.update(extra_num_dict or {})
        self.header.update(extra_str_dict or {})
        with open(os.path.join(self.work_dir, "dglm_params.json"), "w") as f:
            json.dump(self.header, f, indent=1)
        if verbose >= 1:
            print("Save dglm_params.json in WORK_DIR/")

    def link_codes(self, filename_list, verbose=1):
        self.filenames = list(filename_list)

    def check_mem_size(self, class_filename=None, verbose=1):
        """Device bytes held by a solver context for this problem (FP64)."""
        K, Np, nv = self.Nelem, self.Np, self.Nvar
        b = 8 * (3 * nv * K * Np + 25 * K + 4 * K) + 4 * 4 * K + (8 * 3 * K * Np if nv == 1 else 0)
        if verbose >= 1:
            print(f"total bytes = {b / 1e6:.2f} MB")
        return b

    def build(self, filename_list=None, use_sanitize=False, verbose=1):
        """The library is prebuilt (fail loudly if it is missing).  With
        use_sanitize (the reference compiles its CPU solver with
        -fsanitize=address,undefined, dglm_3d_launcher.py:389-390) the host-side
        C++ of the library is built under AddressSanitizer and UBSan and run
        over the operators, generated meshes and this work directory's datain/;
        a sanitizer report or a datain error raises RuntimeError."""
        A.lib()
        if not use_sanitize:
            return
        from . import build as B

        B.build_selfcheck(verbose=verbose >= 2)
        runs = [()]
        if os.path.isdir(getattr(self, "datain_dir", "") or ""):
            runs.append((self.datain_dir, self.N))
        for args in runs:
            rc, out = B.run_selfcheck(*args)
            if verbose >= 1:
                print(out.strip())
            if rc != 0:
                raise RuntimeError(f"sanitized host self-check failed (status {rc}):\n{out}")

    def set_initial_state(self, u0, velocity=None):
        """Override the reference initial condition: u0 [Nvar][K*Np]."""
        self._u0 = np.asarray(u0, dtype=np.float64).reshape(self.Nvar, -1)
        self._velocity = velocity

    def set_dirichlet(self, func):
        """Boundary states of the Dirichlet faces (EtoB == 1; the reference's
        calc_fluxLM leaves them as "ub = function of x,y,z,t",
        euler_dglm.cpp:136-141): func(x, y, z, t) -> [Nvar, n] at the n face
        nodes given.  run() then steps one step at a time and sets the states
        at the time each step starts."""
        self._dirichlet = func

    def _physics(self, filename):
        base = os.path.basename(filename)
        if base.startswith("advec"):
            return "advection"
        if base.startswith("euler"):
            return "euler"
        return "euler" if self.Nvar == 5 else "advection"

    def _write_u(self, u, tstep, physics):
        """dataout/u_NNNNNNN.bin (advection) or u1_NNNNNNN.bin (Euler density)
        as REAL (euler_data.cpp:180-187)."""
        name = ("u_" if physics == "advection" else "u1_") + str(tstep).zfill(7) + ".bin"
        np.ascontiguousarray(u[0], dtype=self.param["dtype"]).tofile(os.path.join(self.dataout_dir, name))

    def run(self, filename, omp_threads=None, verbose=1):
        """Run to TMAX; returns (last_tstep, avg_dt, elapsed_time) like the reference."""
        physics = self._physics(filename)
        # dtype 'f4' (REAL = float in the reference, dglm_3d_launcher.py:138-141,
        # 154; its GPU path then contracts with cublasSgemmBatched): the f4
        # datain files are read and widened, the operator contractions run in
        # FP32 on MFMA (orders 1..6; FP64 above), and dataout is written as f4.
        # The LSRK state, pressure, wave speed and dt stay FP64.
        real = np.float32 if self.param["dtype"] == "f4" else np.float64
        prec = "fp32" if real is np.float32 and self.N <= 6 else "fp64"
        start = time.perf_counter()
        s = Solver(N=self.N, physics=physics, cfl=self.param["CFL"], gamma=self.param["gamma"],
                   tau_scale=self.param["tau_scale"], tmax=float(self.tmax), precision=prec, device=self.devid,
                   datain_dir=self.datain_dir, matrix_order=getattr(self, "matrix_order", "c"))
        px, py, pz = (np.fromfile(os.path.join(self.datain_dir, f"{c}.bin"), dtype=real).astype(np.float64)
                      for c in ("px", "py", "pz"))
        if self._u0 is not None:
            u0 = self._u0
        elif physics == "advection":
            u0 = ic_advection_sine(px, py, pz)
        else:
            u0 = ic_euler_density_wave(px, py, pz, self.param["gamma"])
        if physics == "advection" and self._velocity is not None:
            s.set_velocity(*self._velocity)
        s.set_state(u0)
        self._write_u(u0, 0, physics)
        faces = s.dirichlet_faces()
        if len(faces) and self._dirichlet is None:
            s.close()
            raise ValueError("the mesh has Dirichlet faces (EtoB == 1): give their states with set_dirichlet(func)")
        if len(faces):
            fmask = np.fromfile(os.path.join(self.datain_dir, "Fmask.bin"), dtype=np.int32).reshape(4, -1)
            nodes = ((faces // 4)[:, None] * self.Np + fmask[faces % 4]).ravel()
            fx, fy, fz = px[nodes], py[nodes], pz[nodes]

        def advance(n):
            if not len(faces):
                return s.step(n, return_dt=True)
            out = np.empty(n)
            for i in range(n):
                s.set_dirichlet(np.asarray(self._dirichlet(fx, fy, fz, s.time()[0]), dtype=np.float64))
                out[i] = s.step(1, return_dt=True)[0]
            return out

        # The reference loop (euler_main.cpp:8-15): step; print and write when
        # tstep is a multiple of the cadence or the step was clipped to TMAX.
        # Here the steps run in chunks that end on those multiples, with dt kept
        # on the device; the chunk's dt history gives the printed dt.
        lines = []
        dts = advance(1)
        print_tstep = max(1, find_print_tstep(self.tmax, dts[0], 30))
        while True:
            t, _, ts, done = s.time()
            active = dts[dts > 0]
            if len(active) and (ts % print_tstep == 0 or done):
                line = f"tstep={ts}, dt={active[-1]:.8f}, t={t:.5f}"
                lines.append(line)
                if verbose >= 2:
                    print(line)
                self._write_u(s.get_state(), ts, physics)
            if done:
                break
            dts = advance(print_tstep - ts % print_tstep)
        s.close()
        elapsed = time.perf_counter() - start
        self.stdout = "\n".join(lines) + "\n"
        m1 = re.findall(r"tstep=(\d+)", self.stdout)
        m2 = re.findall(r"dt=([0-9.]+)", self.stdout)
        if not m1:
            return -1, -1, -1
        last_tstep = int(m1[-1])
        avg_dt = np.array(m2, "f8").mean()
        if verbose >= 1:
            print(f"tstep={last_tstep}, dt={avg_dt:.5f}, elapsed_time={elapsed:.5f}")
        return last_tstep, avg_dt, elapsed

    def l2_error(self, ref_sol, numeric_sol, apply_sqrt=True):
        """sqrt(sum_e J_e E^T M_high E), E = u_ref(high nodes) - interpM u_e (:328-355)."""
        K, Np, hNp = self.Nelem, self.Np, self.tet.high_Np
        u = np.asarray(numeric_sol, dtype=np.float64).reshape(K, Np)
        r = np.asarray(ref_sol, dtype=np.float64).reshape(K, hNp)
        E = r - u @ self.op_ref.interpM.T
        Eint = self.mesh.J * np.einsum("ei,ij,ej->e", E, self.op_ref.high_M, E)
        s = np.sum(Eint)
        return np.sqrt(s) if apply_sqrt else s


class DGLM3DSingleCPU(DGLM3DBase):
    def init_mesh(self, mesh_info, verbose=1):
        if mesh_info["type"] == "gmsh":
            from .gmsh import Gmsh3DSingle

            elem = Gmsh3DSingle(mesh_info["file"])  # *.msh file generated by Gmsh (ASCII 2.2 or 4.1)
            if verbose >= 1:
                print(f"Init Gmsh file: {mesh_info['file']}")
        elif mesh_info["type"] == "uniform":
            elem = Uniform3DSingle(*mesh_info["param"])
            if verbose >= 1:
                print(f"Init uniform mesh: {mesh_info['param']}")
        else:
            raise ValueError(f"mesh type must be 'gmsh' or 'uniform', got {mesh_info['type']!r}")
        self.mesh = Mesh3D(self.op_face, elem)
        self.Nelem = self.mesh.Nelem


class DGLM3DSingleGPU(DGLM3DSingleCPU):
    def __init__(self, dglm_dir, work_dir, N, Nvar, tmax, dglm_param, devid=0):
        super().__init__(dglm_dir, work_dir, N, Nvar, tmax, dglm_param)
        self.devid = devid
        self.pNP = ((self.Np - 1) // 32 + 1) * 32
        self.pNFF = ((self.Nface * self.Nfp - 1) // 32 + 1) * 32
