"""Parity of the HIP path (through the C ABI) with the CPU oracle.

Bounds: the FP64 kernels sum the contraction in a different order from the
oracle (and use FMA), so single RHS evaluations agree to ~1e-15 relative and
runs of 50-200 steps to ~1e-13; the required bound is 1e-12 relative L2 per
conserved variable."""
import math
import os
import sys
import tempfile

import numpy as np
import pytest

from helpers import noisy_state, rel_err, rel_err_per_var, smooth_state
from exadglm_amd.problem import (cube_problem, ic_advection_sine, ic_euler_density_wave, isentropic_vortex, l2_error,
                                 uniform_problem, vortex_box)
from exadglm_amd.solver import Solver
import pyoracle as O

pytestmark = pytest.mark.gpu
PHYS = ["advection", "euler"]


def _say(msg):
    print(msg, flush=True)


def euler_errors(u, uo):
    """Relative L2 per conserved variable; rho*w, which is zero in the exact
    z-extruded vortex (its own norm is discretisation noise), is measured
    against the norm of the whole momentum."""
    errs = rel_err_per_var(u, uo)
    errs[3] = float(np.linalg.norm(u[3] - uo[3]) / np.linalg.norm(uo[1:4]))
    return errs


def _variants(N, physics):
    """Stage kernels built for this order and physics."""
    return ["simple"] + (["mfma"] if N <= 6 else []) + (["mfma_ws"] if physics == "euler" and N in (3, 4) else [])


@pytest.mark.parametrize("physics", PHYS)
@pytest.mark.parametrize("N,variant", [(n, v) for n in range(1, 7) for v in ("simple", "mfma", "mfma_ws")])
def test_rhs_parity(gpu, physics, N, variant):
    if variant not in _variants(N, physics):
        pytest.skip(f"no {variant} kernel for {physics} N={N}")
    _say(f"rhs parity {physics} N={N} {variant}")
    pb = cube_problem(N, 3)
    u = smooth_state(pb, physics, seed=N + 11)
    with Solver(pb, physics) as s:
        s.set_stage_kernel(variant)
        assert s.stage_kernel() == variant
        r, mv, tau = s.rhs(u)
    ro, mvo, tauo = O.Oracle(pb, physics).rhs(u)
    assert rel_err(mv, mvo) < 1e-15
    assert rel_err(tau, tauo) < 1e-15
    # The RHS of a smooth field is a small difference of volume and surface
    # terms; bound the pointwise error by the term scale
    # (|operator row| x |metric| x |flux|) and the relative L2 by 1e-12.
    a = pb.arrays
    opn = max(np.abs(a[k]).sum(axis=1).max() for k in ("WDr", "WDs", "WDt", "LIFT"))
    met = max(np.abs(a[k]).max() for k in ("rx", "ry", "rz", "sx", "sy", "sz", "tx", "ty", "tz", "Fscale"))
    for v, e in enumerate(rel_err_per_var(r, ro)):
        assert e < 1e-12
        assert np.abs(r[v] - ro[v]).max() < 1e-15 * 3 * opn * met * np.abs(u).max() ** 2


@pytest.mark.parametrize("physics", PHYS)
@pytest.mark.parametrize("N", [7, 8, 9, 10])
def test_high_order_parity(gpu, physics, N):
    """Orders 7-10 run on the LDS-staged FP64 kernel (no MFMA kernel is
    offered): one RHS, with its wave speeds and tau, and 3 LSRK steps against
    the oracle on 2^3 hexes x 6 = 48 tets."""
    _say(f"high-order parity {physics} N={N}")
    pb = cube_problem(N, 2)
    u = smooth_state(pb, physics, seed=N + 31)
    with Solver(pb, physics, cfl=0.5) as s:
        assert s.stage_kernel() == "simple"
        with pytest.raises(RuntimeError):
            s.set_stage_kernel("mfma")
        r, mv, tau = s.rhs(u)
        s.set_state(u)
        s.step(3)
        u3 = s.get_state()
    o = O.Oracle(pb, physics, cfl=0.5)
    ro, mvo, tauo = o.rhs(u)
    uo3 = o.run(u, 3)[0]
    assert rel_err(mv, mvo) < 1e-15
    assert rel_err(tau, tauo) < 1e-15
    assert max(rel_err_per_var(r, ro)) < 1e-12
    assert max(rel_err_per_var(u3, uo3)) < 1e-12


def oracle_run_chunked(o, u0, nsteps, capsys, label, chunk=20):
    """o.run in chunks (LSRK4(5) starts every step with a = 0, so the register
    k does not carry over), printing progress past pytest's capture so that a
    long CPU run is visibly alive."""
    u, t, dts = np.array(u0), 0.0, []
    for n0 in range(0, nsteps, chunk):
        u, d, t = o.run(u, min(chunk, nsteps - n0), t=t)
        dts.append(d)
        with capsys.disabled():
            print(f"  {label}: oracle {n0 + len(d)}/{nsteps} steps", flush=True)
    return u, np.concatenate(dts), t


@pytest.mark.parametrize("flux", ["lm", "rusanov"])
def test_c1_advection_100_steps(gpu, flux):
    """Linear advection N=3 on 10^3 hexes x 6 = 6000 tets, v=(1,1,1),
    sin sin sin, CFL 0.5, 100 steps, with the configuration's upwind flux
    (flux option 'rusanov' = exact upwinding for advection) and with the
    reference's LM flux."""
    _say(f"C1 advection K=6000 N=3 100 steps ({flux})")
    pb = cube_problem(3, 10)
    assert pb.K == 6000
    u0 = ic_advection_sine(pb.arrays["px"], pb.arrays["py"], pb.arrays["pz"])
    with Solver(pb, "advection", cfl=0.5, flux=flux) as s:
        s.set_state(u0)
        dts = s.step(100, return_dt=True)
        u = s.get_state()
        t, _, ts, _ = s.time()
    uo, dto, to = O.Oracle(pb, "advection", cfl=0.5, flux=flux).run(u0, 100)
    assert ts == 100
    assert rel_err(dts, dto) < 1e-14
    assert abs(t - to) < 1e-13 * to
    assert rel_err(u, uo) < 1e-12


@pytest.mark.parametrize("N", [3, 4, 5, 6])
def test_euler_vortex_steps(gpu, N):
    _say(f"Euler vortex N={N} 6^3 hexes 40 steps")
    pb = uniform_problem(N, 7, 7, 7, box=vortex_box(6))
    a = pb.arrays
    u0 = isentropic_vortex(a["px"], a["py"], a["pz"])
    with Solver(pb, "euler", cfl=0.4) as s:
        s.set_state(u0)
        dts = s.step(40, return_dt=True)
        u = s.get_state()
    uo, dto, _ = O.Oracle(pb, "euler", cfl=0.4).run(u0, 40)
    assert rel_err(dts, dto) < 1e-13
    errs = euler_errors(u, uo)
    assert max(errs) < 1e-12, errs


def test_advection_velocity_field(gpu):
    """Advection with v = (1, 0.5, 0.25), N=5."""
    _say("advection v=(1,.5,.25) N=5")
    pb = cube_problem(5, 4)
    u0 = smooth_state(pb, "advection", seed=4)
    v = (1.0, 0.5, 0.25)
    with Solver(pb, "advection", cfl=0.5) as s:
        s.set_velocity(*v)
        s.set_state(u0)
        s.step(30)
        u = s.get_state()
    uo, _, _ = O.Oracle(pb, "advection", cfl=0.5, velocity=v).run(u0, 30)
    assert rel_err(u, uo) < 1e-12


@pytest.mark.timeout(900)
@pytest.mark.parametrize("flux", ["rusanov", "lm"])
def test_c2_full_size_200_steps(gpu, flux, capsys):
    """The benchmark workload at full size: Euler N=4 on 32^3 hexes x 6 =
    196,608 tets, isentropic vortex, CFL 0.4, 200 steps with the
    configuration's Rusanov flux and with the reference's LM flux, against
    the oracle: relative L2 <= 1e-12 per conserved variable, dt history
    <= 1e-13.  (The oracle takes one to two minutes per flux on the host.)"""
    _say(f"C2 full size K=196608 N=4 200 steps ({flux})")
    pb = uniform_problem(4, 33, 33, 33, box=vortex_box(32))
    assert pb.K == 196608
    a = pb.arrays
    u0 = isentropic_vortex(a["px"], a["py"], a["pz"])
    with Solver(pb, "euler", cfl=0.4, flux=flux) as s:
        assert s.stage_kernel() == "mfma_ws"
        s.set_state(u0)
        dts = s.step(200, return_dt=True)
        u = s.get_state()
    uo, dto, _ = oracle_run_chunked(O.Oracle(pb, "euler", cfl=0.4, flux=flux), u0, 200, capsys, f"C2 {flux}")
    assert rel_err(dts, dto) < 1e-13
    errs = euler_errors(u, uo)
    _say(f"  per-variable rel L2 vs oracle: {errs}")
    assert max(errs) < 1e-12, errs


def test_c2_full_size_conservation_and_determinism(gpu):
    """C2 at full size: mass, momentum and energy conserved over 20 steps
    (graph replay, Rusanov flux), and a rerun is bitwise the same."""
    _say("C2 full size conservation + determinism")
    pb = uniform_problem(4, 33, 33, 33, box=vortex_box(32))
    a = pb.arrays
    u0 = isentropic_vortex(a["px"], a["py"], a["pz"])
    with Solver(pb, "euler", cfl=0.4, flux="rusanov") as s:
        s.set_state(u0)
        s.step(3)
        I0 = s.integrals(pb.mesh.J, pb.op_ref.M)
        s.step(20)
        I1 = s.integrals(pb.mesh.J, pb.op_ref.M)
        u23 = s.get_state()
        s.set_state(u0)
        s.step(23)
        u23b = s.get_state()
    for v in (0, 1, 2, 4):
        assert abs(I1[v] - I0[v]) <= 1e-12 * abs(I0[v])
    assert np.array_equal(u23, u23b)


def _adv_exact(pb, t):
    m = pb.mesh
    x, y, z = m.high_PX.ravel(), m.high_PY.ravel(), m.high_PZ.ravel()
    xv, yv, zv = ((x - t) + 1) % 1 - 1, ((y - t) + 1) % 1 - 1, ((z - t) + 1) % 1 - 1
    return np.sin(2 * np.pi * xv) * np.sin(2 * np.pi * yv) * np.sin(2 * np.pi * zv)


def test_published_l2_advection(gpu):
    """ex1.advection.11.cpu.single.ipynb:214: 809 steps, L2 = 2.284502600150407e-05."""
    _say("published advection L2")
    pb = cube_problem(5, 6)
    u0 = ic_advection_sine(pb.arrays["px"], pb.arrays["py"], pb.arrays["pz"])
    with Solver(pb, "advection", cfl=0.8, tmax=2.0) as s:
        s.set_state(u0)
        dts = s.step(850, return_dt=True)
        t, _, ts, done = s.time()
        u = s.get_state()
    assert ts == 809 and done and t == 2.0 and int((dts > 0).sum()) == 809
    l2 = l2_error(pb, _adv_exact(pb, 2.0), u[0])
    assert abs(l2 - 2.284502600150407e-05) / 2.284502600150407e-05 < 1e-9


def test_published_l2_euler(gpu):
    """ex3.euler.11.cpu.single.ipynb:221: 1318 steps, L2(rho) = 2.429869462976010e-05."""
    _say("published Euler L2")
    pb = cube_problem(5, 6)
    u0 = ic_euler_density_wave(pb.arrays["px"], pb.arrays["py"], pb.arrays["pz"])
    with Solver(pb, "euler", cfl=0.8, tmax=2.0) as s:
        s.set_state(u0)
        s.step(1400)
        t, _, ts, done = s.time()
        u = s.get_state()
    assert ts == 1318 and done
    m = pb.mesh
    x, y, z = m.high_PX.ravel(), m.high_PY.ravel(), m.high_PZ.ravel()
    xv, yv, zv = ((x - 2) + 1) % 1 - 1, ((y + 1) + 1) % 1 - 1, ((z - 2) + 1) % 1 - 1
    l2 = l2_error(pb, 1 + 0.2 * np.sin(2 * np.pi * (xv + yv + zv)), u[0])
    assert abs(l2 - 2.429869462976010e-05) / 2.429869462976010e-05 < 1e-9


def test_create_from_datain_dir(gpu):
    """Drop-in on a work dir written in the reference's datain format, both
    matrix orders ('c' for the CPU build, 'f' for the cuBLAS build)."""
    _say("datain dir drop-in")
    pb = cube_problem(4, 3)
    u0 = smooth_state(pb, "euler", seed=5)
    with Solver(pb, "euler", cfl=0.5) as s:
        s.set_state(u0)
        s.step(5)
        ua = s.get_state()
    for order in ("c", "f"):
        with tempfile.TemporaryDirectory() as d:
            O.write_datain(pb, d, matrix_order=order)
            with Solver(N=4, physics="euler", cfl=0.5, datain_dir=d, matrix_order=order) as s:
                s.set_state(u0)
                s.step(5)
                ub = s.get_state()
        assert np.array_equal(ua, ub)


def test_timing_and_noop_after_tmax(gpu):
    _say("timing + tmax no-op")
    pb = cube_problem(3, 3)
    u0 = smooth_state(pb, "euler", seed=2)
    with Solver(pb, "euler", cfl=0.5, tmax=0.01) as s:
        s.set_state(u0)
        s.set_timing(True)
        s.step(200)
        tm = s.timing()
        t, _, ts, done = s.time()
        u = s.get_state()
        s.step(3)
        assert np.array_equal(u, s.get_state())
    assert done and t == 0.01
    assert tm["stage"][1] == 1000 and tm["stage"][0] > 0
    uo, _, _ = O.Oracle(pb, "euler", cfl=0.5, tmax=0.01).run(u0, 200)
    assert rel_err(u, uo) < 1e-12


@pytest.mark.parametrize("physics", PHYS)
@pytest.mark.parametrize("N", [2, 3, 4, 5])
def test_steps_both_stage_kernels(gpu, physics, N):
    """Every stage kernel built for the order tracks the oracle over 20 steps
    (K = 162, a partial last batch)."""
    _say(f"both kernels {physics} N={N}")
    pb = cube_problem(N, 3)
    u0 = smooth_state(pb, physics, seed=21 + N)
    out = {}
    for variant in _variants(N, physics):
        with Solver(pb, physics, cfl=0.5) as s:
            s.set_stage_kernel(variant)
            s.set_state(u0)
            s.step(20)
            out[variant] = s.get_state()
    uo, _, _ = O.Oracle(pb, physics, cfl=0.5).run(u0, 20)
    for variant, u in out.items():
        for e in rel_err_per_var(u, uo):
            assert e < 1e-12, variant


@pytest.mark.parametrize("mesh", ["uniform", "gmsh"])
def test_launcher_drop_in(gpu, tmp_path, mesh):
    """The notebook flow of ex1.advection.21.gpu.single.ipynb through the
    launcher-compatible API: 809 steps and the published L2; with the same
    periodic box also read from a Gmsh .msh file (MSH 4.1 with its periodic
    node pairs)."""
    _say(f"launcher drop-in ({mesh} mesh)")
    from exadglm_amd.launcher import DGLM3DSingleGPU

    dg = DGLM3DSingleGPU(None, str(tmp_path / "ex1"), 5, 1, 2, {"CFL": 0.8, "tau_scale": 1.2}, devid=0)
    param = [7, 7, 7, 0, 1, 0, 1, 0, 1, {"x": "pbc", "y": "pbc", "z": "pbc"}]
    if mesh == "uniform":
        dg.init_mesh({"type": "uniform", "param": param}, verbose=0)
    else:
        from exadglm_amd.mesh import Uniform3DSingle
        from test_gmsh import _periodic_pairs, write_msh

        u = Uniform3DSingle(*param)
        path = str(tmp_path / "box.msh")
        write_msh(path, u.VXYZ, u.EtoV, (), _periodic_pairs(u.VXYZ, (0, 1, 0, 1, 0, 1), [0, 1, 2]), "4.1")
        dg.init_mesh({"type": "gmsh", "file": path}, verbose=0)
    dg.make_work_dir()
    dg.save_to_binaries(matrix_order="f", verbose=0)
    dg.save_to_header(["advec_nopad_data_host.cu"], verbose=0)
    dg.link_codes(["advec_nopad_main.cu"], verbose=0)
    dg.build(["advec_nopad_main.cu"], verbose=0)
    tstep, avg_dt, elapsed = dg.run("advec_nopad_main", verbose=0)
    assert tstep == 809
    import glob
    files = sorted(glob.glob(dg.dataout_dir + "/*.bin"))
    assert files[0].endswith("u_0000000.bin") and files[-1].endswith("u_0000809.bin")
    u = np.fromfile(files[-1])
    m = dg.mesh
    x, y, z = m.high_PX.ravel(), m.high_PY.ravel(), m.high_PZ.ravel()
    xv, yv, zv = ((x - 2) + 1) % 1 - 1, ((y - 2) + 1) % 1 - 1, ((z - 2) + 1) % 1 - 1
    ref = np.sin(2 * np.pi * xv) * np.sin(2 * np.pi * yv) * np.sin(2 * np.pi * zv)
    l2 = dg.l2_error(ref, u)
    assert abs(l2 - 2.284502600150407e-05) / 2.284502600150407e-05 < 1e-9


@pytest.mark.parametrize("dtype", ["f8", "f4"])
def test_launcher_euler_notebook_flow(gpu, tmp_path, dtype):
    """The notebook flow of ex3.euler.21.gpu.single.ipynb (density wave,
    N=5, 6^3 hexes x 6 = 1296 tets, CFL 0.8, tmax 2) through the
    launcher-compatible API.  dtype 'f8': 1318 steps and the published L2
    error 2.429869462976010e-05 (ex3.euler.11.cpu.single.ipynb:221).  dtype
    'f4' (the reference's REAL = float mode): f4 datain and dataout files, the
    FP32-MFMA contraction; its final density equals the FP64 oracle run on the
    same f4-rounded datain within the f4 rounding of the output (relative L2
    3e-7: the f4 output rounding is 3e-8 RMS, and the FP32 contraction of
    the interpolation remainders on this coarse mesh (the wave spans ~8
    elements) adds ~1e-7 over the 1318 steps)."""
    _say(f"launcher Euler notebook flow ({dtype})")
    from glob import glob

    from exadglm_amd.launcher import DGLM3DSingleGPU

    dg = DGLM3DSingleGPU(None, str(tmp_path / "ex3"), 5, 5, 2,
                         {"CFL": 0.8, "gamma": 1.4, "tau_scale": 1.2, "interpN": 2, "dtype": dtype}, devid=0)
    dg.init_mesh({"type": "uniform", "param": [7, 7, 7, 0, 1, 0, 1, 0, 1, {"x": "pbc", "y": "pbc", "z": "pbc"}]},
                 verbose=0)
    dg.make_work_dir()
    dg.save_to_binaries(matrix_order="f", verbose=0)
    names = [f"euler_nopad_{n}.cu" for n in ("data_host", "data_dev", "dglm", "time", "init", "main")]
    dg.save_to_header([n for n in names if "main" not in n], verbose=0)
    dg.link_codes(names, verbose=0)
    dg.build([n for n in names if "main" in n], verbose=0)
    tstep, avg_dt, _ = dg.run("euler_nopad_main", verbose=0)
    real = np.float32 if dtype == "f4" else np.float64
    files = sorted(glob(dg.dataout_dir + "/*.bin"))
    assert files[0].endswith("u1_0000000.bin") and files[-1].endswith(f"u1_{tstep:07d}.bin")
    u = np.fromfile(files[-1], dtype=real)
    assert u.size == dg.Nelem * dg.Np
    m = dg.mesh
    x, y, z = m.high_PX.ravel(), m.high_PY.ravel(), m.high_PZ.ravel()
    wrap = lambda q: (q + 1) % 1 - 1  # noqa: E731
    ref = 1 + 0.2 * np.sin(2 * np.pi * (wrap(x - 2) + wrap(y + 0.5 * 2) + wrap(z - 2)))
    l2 = dg.l2_error(ref, u)
    _say(f"  {tstep} steps, avg dt {avg_dt:.6f}, L2 {l2:.15e}")
    if dtype == "f8":
        assert tstep == 1318
        assert abs(l2 - 2.429869462976010e-05) / 2.429869462976010e-05 < 1e-9
        return
    # the FP64 oracle on the f4 datain, widened exactly as the library reads it
    d = dg.datain_dir
    rd = lambda n, dt=real: np.fromfile(os.path.join(d, n + ".bin"), dtype=dt)  # noqa: E731
    arrays = {n: rd(n, np.int32) for n in ("vmapP", "vmapF", "Fmask", "EtoB")}
    for n in ("Fscale", "rx", "ry", "rz", "sx", "sy", "sz", "tx", "ty", "tz", "nx", "ny", "nz", "px", "py", "pz"):
        arrays[n] = rd(n).astype(np.float64)
    for n in ("WDr", "WDs", "WDt", "LIFT"):
        arrays[n] = np.ascontiguousarray(rd(n).astype(np.float64).reshape((dg.Np, -1), order="F"))

    class _P:
        pass

    pb = _P()
    pb.arrays, pb.K, pb.N, pb.Np, pb.Nfp = arrays, dg.Nelem, dg.N, dg.Np, dg.Nfp
    u0 = ic_euler_density_wave(arrays["px"], arrays["py"], arrays["pz"])
    uo, dto, to = O.Oracle(pb, "euler", cfl=0.8, tau_scale=1.2, tmax=2.0).run(u0, tstep)
    assert abs(to - 2.0) < 1e-12 and (dto > 0).sum() == tstep
    e = rel_err(u.astype(np.float64), uo[0])
    _say(f"  f4 density vs FP64 oracle on the f4 datain: {e:.3e}")
    assert e < 3e-7


MIXED = ["mfma_mx", "mfma"]  # the fused mixed-precision kernel (AUTO) and the generic one


@pytest.mark.parametrize("kernel", MIXED)
@pytest.mark.parametrize("flux", ["lm", "rusanov"])
@pytest.mark.parametrize("N", [1, 2, 3, 4, 5, 6])
def test_mixed_precision_rhs(gpu, N, flux, kernel):
    """FP32-MFMA contraction of the split fluxes (the remainders of the
    least-squares affine fit of every flux block, in the fused and in the
    generic kernel): one RHS within FP32 rounding of the in-block flux
    variation (bound 1e-7 of the term scale), both fluxes."""
    _say(f"mixed precision rhs N={N} {flux} {kernel}")
    pb = cube_problem(N, 3)
    u = smooth_state(pb, "euler", seed=N + 31)
    with Solver(pb, "euler", precision="fp32", flux=flux) as s:
        # AUTO: the FP64 warp-specialised kernel where it exists (faster there)
        assert s.stage_kernel() == ("mfma_ws" if N in (3, 4) else "mfma_mx")
        s.set_stage_kernel(kernel)
        assert s.stage_kernel() == kernel
        r, mv, tau = s.rhs(u)
    ro, mvo, tauo = O.Oracle(pb, "euler", flux=flux).rhs(u)
    assert rel_err(mv, mvo) < 1e-15  # wave speed stays FP64
    a = pb.arrays
    opn = max(np.abs(a[k]).sum(axis=1).max() for k in ("WDr", "WDs", "WDt", "LIFT"))
    met = max(np.abs(a[k]).max() for k in ("rx", "ry", "rz", "sx", "sy", "sz", "tx", "ty", "tz", "Fscale"))
    for v in range(5):
        assert np.abs(r[v] - ro[v]).max() < 1e-7 * 3 * opn * met * np.abs(u).max() ** 2


@pytest.mark.parametrize("kernel", MIXED)
@pytest.mark.parametrize("N", [4, 5, 6])
def test_mixed_precision_vortex(gpu, N, kernel):
    """Isentropic vortex, 6^3 hexes, 40 steps: FP32-MFMA contraction with FP64
    state / pressure / wave speed stays within 1e-8 relative L2 of the FP64
    oracle for rho, rho*u, rho*v and E; rho*w (zero in the exact solution) is
    measured against the momentum norm."""
    _say(f"mixed precision vortex N={N} {kernel}")
    pb = uniform_problem(N, 7, 7, 7, box=vortex_box(6))
    a = pb.arrays
    u0 = isentropic_vortex(a["px"], a["py"], a["pz"])
    with Solver(pb, "euler", cfl=0.4, precision="fp32") as s:
        s.set_stage_kernel(kernel)
        s.set_state(u0)
        s.step(40)
        u = s.get_state()
    uo, _, _ = O.Oracle(pb, "euler", cfl=0.4).run(u0, 40)
    errs = euler_errors(u, uo)
    _say(f"  per-variable rel L2 vs oracle: {errs}")
    assert max(errs) < 1e-8, errs


def _term_scale(pb, u):
    """Scale of the RHS terms: operator row sums x metric x |u|^2."""
    a = pb.arrays
    opn = max(np.abs(a[k]).sum(axis=1).max() for k in ("WDr", "WDs", "WDt", "LIFT"))
    met = max(np.abs(a[k]).max() for k in ("rx", "ry", "rz", "sx", "sy", "sz", "tx", "ty", "tz", "Fscale"))
    return 3 * opn * met * np.abs(u).max() ** 2


# split-BF16 contraction: (relative bound of one RHS against the term scale,
# bound of 40 vortex steps against the oracle).  Three terms keep ~16 bits of
# each operand (CPU emulation of the scheme at N=6, 6^3 hexes, 40 steps:
# 3.9e-7); six terms ~24 bits (emulation with one FP32 accumulator 4.7e-9;
# the kernel sums the cross terms in a second accumulator).
BF16 = {"bf16x3": (3e-7, 1e-6), "bf16x6": (1e-8, 1e-8)}


@pytest.mark.parametrize("prec", sorted(BF16))
@pytest.mark.parametrize("flux", ["lm", "rusanov"])
@pytest.mark.parametrize("N", [1, 2, 3, 4, 5, 6])
def test_split_bf16_rhs(gpu, N, flux, prec):
    """Split-BF16 contraction (bf16 MFMA on the least-squares affine-fit
    remainders, FP32 accumulation, the fit's term in FP64): one RHS against
    the FP64 oracle, both fluxes; the wave speed stays FP64."""
    _say(f"split-bf16 rhs N={N} {flux} {prec}")
    pb = cube_problem(N, 3)
    u = smooth_state(pb, "euler", seed=N + 37)
    with Solver(pb, "euler", precision=prec, flux=flux) as s:
        assert s.stage_kernel() == "mfma"
        r, mv, tau = s.rhs(u)
    ro, mvo, tauo = O.Oracle(pb, "euler", flux=flux).rhs(u)
    assert rel_err(mv, mvo) < 1e-15
    scale = _term_scale(pb, u)
    err = max(np.abs(r[v] - ro[v]).max() for v in range(5)) / scale
    _say(f"  max error / term scale {err:.2e}")
    assert err < BF16[prec][0]


@pytest.mark.parametrize("prec", sorted(BF16))
def test_split_bf16_advection_rhs(gpu, prec):
    """The split-BF16 contraction for advection (N=4, LM flux)."""
    pb = cube_problem(4, 3)
    u = smooth_state(pb, "advection", seed=5)
    with Solver(pb, "advection", precision=prec) as s:
        r, _, _ = s.rhs(u)
    ro, _, _ = O.Oracle(pb, "advection").rhs(u)
    # linear flux, default velocity (1, 1, 1): terms scale with |u|, not |u|^2
    scale = _term_scale(pb, u) / np.abs(u).max()
    err = np.abs(r - ro).max() / scale
    _say(f"  advection split-bf16 {prec}: max error / term scale {err:.2e}")
    assert err < BF16[prec][0]


@pytest.mark.parametrize("prec", sorted(BF16))
@pytest.mark.parametrize("N", [4, 6])
def test_split_bf16_vortex(gpu, N, prec):
    """Isentropic vortex, 6^3 hexes, 40 steps with the split-BF16
    contraction against the FP64 oracle (per-variable relative L2)."""
    _say(f"split-bf16 vortex N={N} {prec}")
    pb = uniform_problem(N, 7, 7, 7, box=vortex_box(6))
    a = pb.arrays
    u0 = isentropic_vortex(a["px"], a["py"], a["pz"])
    with Solver(pb, "euler", cfl=0.4, precision=prec) as s:
        s.set_state(u0)
        s.step(40)
        u = s.get_state()
    uo, _, _ = O.Oracle(pb, "euler", cfl=0.4).run(u0, 40)
    errs = euler_errors(u, uo)
    _say(f"  per-variable rel L2 vs oracle: {errs}")
    assert max(errs) < BF16[prec][1], errs


@pytest.mark.timeout(600)
def test_c3_split_bf16_full_size(gpu):
    """The mixed-precision configuration at full size (N=6, 82,944 tets,
    Rusanov, 500 steps) with the six-term split-BF16 contraction: within
    1e-8 relative L2 per conserved variable of the FP64 run (itself pinned
    against the oracle at this size in test_c3_mixed_precision_full_size)."""
    _say("C3 full size K=82944 N=6 500 steps (split-BF16 x6 vs FP64)")
    pb = uniform_problem(6, 25, 25, 25, box=vortex_box(24))
    a = pb.arrays
    u0 = isentropic_vortex(a["px"], a["py"], a["pz"])
    out = {}
    for prec in ("fp64", "bf16x6"):
        with Solver(pb, "euler", cfl=0.4, flux="rusanov", precision=prec) as s:
            s.set_state(u0)
            s.step(500)
            out[prec] = s.get_state()
    errs = euler_errors(out["bf16x6"], out["fp64"])
    _say(f"  split-BF16 x6 vs FP64 after 500 steps: {errs}")
    assert max(errs) < 1e-8, errs


@pytest.mark.timeout(600)
def test_c3_mixed_precision_full_size(gpu, capsys):
    """The mixed-precision configuration at full size: isentropic vortex,
    N=6 on 24^3 hexes x 6 = 82,944 tets, Rusanov flux, CFL 0.4, 500 steps.
    The FP64 GPU run is pinned against the oracle over its first 5 steps at
    this size (<= 1e-12 per variable; over 40 steps on 6^3 hexes elsewhere);
    the mixed-precision run (FP32-MFMA contraction) stays within 1e-8
    relative L2 per conserved variable of the FP64 run after 500 steps, and
    both are equally close to the analytic translated vortex."""
    _say("C3 full size K=82944 N=6 500 steps (mixed vs FP64)")
    pb = uniform_problem(6, 25, 25, 25, box=vortex_box(24))
    assert pb.K == 82944
    a = pb.arrays
    u0 = isentropic_vortex(a["px"], a["py"], a["pz"])
    out = {}
    for prec in ("fp64", "fp32"):
        with Solver(pb, "euler", cfl=0.4, flux="rusanov", precision=prec) as s:
            assert s.stage_kernel() == ("mfma_mx" if prec == "fp32" else "mfma")
            s.set_state(u0)
            if prec == "fp64":
                s.step(5)
                out["u5"] = s.get_state()
                s.step(495)
            else:
                s.step(500)
            out[prec] = s.get_state()
            t = s.time()[0]
    uo5, _, _ = oracle_run_chunked(O.Oracle(pb, "euler", cfl=0.4, flux="rusanov"), u0, 5, capsys, "C3", chunk=1)
    e5 = euler_errors(out["u5"], uo5)
    assert max(e5) < 1e-12, e5
    errs = euler_errors(out["fp32"], out["fp64"])
    exact = isentropic_vortex(a["px"], a["py"], a["pz"], t=t)
    e64, e32 = euler_errors(out["fp64"], exact), euler_errors(out["fp32"], exact)
    _say(f"  mixed vs FP64 {errs}; vs analytic FP64 {e64}, mixed {e32}")
    assert max(errs) < 1e-8, errs
    for v in range(5):
        assert abs(e32[v] - e64[v]) <= 0.05 * e64[v]


def _vortex_error(N, nhex, T, precision="fp64"):
    """Per-variable nodal relative L2 error against the analytic translated
    vortex after running the vortex workload to time T (Rusanov flux, CFL
    0.4); rho*w against the momentum norm."""
    pb = uniform_problem(N, nhex + 1, nhex + 1, nhex + 1, box=vortex_box(nhex))
    a = pb.arrays
    u0 = isentropic_vortex(a["px"], a["py"], a["pz"])
    with Solver(pb, "euler", cfl=0.4, flux="rusanov", tmax=T, precision=precision) as s:
        s.set_state(u0)
        while not s.time()[3]:
            s.step(50)
        u = s.get_state()
        t = s.time()[0]
    assert abs(t - T) < 1e-12
    exact = isentropic_vortex(a["px"], a["py"], a["pz"], t=T)
    return euler_errors(u, exact)


def test_isentropic_vortex_against_analytic(gpu):
    """The isentropic-vortex initial condition and the FP64 solver together,
    pinned by the analytic translated vortex (no reference counterpart):
    N = 4 on 12^3 and 24^3 hexes of the vortex box, to T = 1 (the vortex
    moves by (1, 1)).  Halving h must cut every variable's error by at least
    2^3.5 (order N+1 = 5 expected; the density error falls 32x in practice),
    and at 24^3 the density error is below 3e-5 (the superposed periodic
    images of the analytic solution are accurate to ~1e-9 there)."""
    _say("isentropic vortex vs analytic, N=4, 12^3 and 24^3 hexes, T=1")
    e12 = _vortex_error(4, 12, 1.0)
    e24 = _vortex_error(4, 24, 1.0)
    rates = [math.log2(e12[v] / e24[v]) for v in range(5)]
    _say(f"  12^3 {e12}\n  24^3 {e24}\n  observed orders {rates}")
    for v in (0, 1, 2, 4):  # rho*w is zero in the exact solution (its error is noise)
        assert rates[v] > 3.5, (v, rates)
    assert e24[0] < 3e-5


@pytest.mark.parametrize("N", [3, 4])
def test_warp_specialised_kernel_many_batches(gpu, N):
    """The persistent warp-specialised kernel with more batches than
    workgroups (XCD-grouped schedule, 500 batches, last one holding 2
    elements) on a discontinuous state (tau term active): RHS and 3 steps
    against the oracle."""
    _say(f"warp-specialised many batches N={N}")
    pb = cube_problem(N, 11)
    assert (pb.K + 15) // 16 == 500 and pb.K % 16 == 2
    u0 = noisy_state(pb, "euler", seed=40 + N)
    o = O.Oracle(pb, "euler", cfl=0.4)
    with Solver(pb, "euler", cfl=0.4) as s:
        s.set_stage_kernel("mfma_ws")
        r, _, _ = s.rhs(u0)
        s.set_state(u0)
        s.step(3)
        u = s.get_state()
    ro, _, _ = o.rhs(u0)
    for e in rel_err_per_var(r, ro):
        assert e < 1e-12
    uo, _, _ = o.run(u0, 3)
    for e in rel_err_per_var(u, uo):
        assert e < 1e-12
