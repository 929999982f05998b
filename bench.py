"""Benchmark: DOF-stage updates/s of the explicit DGLM RHS + LSRK4(5) step.

Headline workload (one GPU): compressible Euler, isentropic vortex (gamma 1.4,
strength 5, centre (5,5), free stream (1,1,1,0,1)) on [0,10]^2 x [0,10/3] with
32^3 hexes x 6 tets = 196,608 tetrahedra, order N=4 (Np=35), Rusanov flux,
LSRK4(5) at CFL 0.4, 200 steps, FP64.  value = K * Np * 5 vars * 5 stages *
steps / wall time, state resident in HBM (mesh, operators and initial state
uploaded before the timed region, the step loop replayed as hipGraphs).

The same JSON line carries the roofline of the dominant kernel (the fused
stage kernel, launch durations from HIP events on the solver's stream), a CPU
baseline (the C restatement of the reference CPU path on this host's cores,
bounded sample) and, unless --no-extra, the secondary figures of the metric:
the same workload with the reference's LM flux; FP64 vs mixed precision on the
N=6 vortex (24^3 hexes, 500 steps) with per-variable error against the FP64
run and the analytic vortex; and the adjoint/forward time ratio of the
Euler adjoint (vortex + Gaussian bumps, 300 steps, ~sqrt(n) checkpoints).
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0      # MI355X HBM3E
FP64_PEAK_TFLOPS = 78.6    # MI355X dense FP64 (vector and matrix)
FP32_PEAK_TFLOPS = 157.3   # MI355X dense FP32 (vector and matrix)


def stage_bytes(K, Np, Nfp, nvar):
    """Compulsory HBM bytes of one fused stage launch: u read, k read, k write,
    u_out write (FP64), the per-element geometry record (25 doubles), packed
    face connectivity (4 int32) and the element's own face wave speeds
    (4 doubles)."""
    return K * (32 * nvar * Np + 8 * 25 + 16 + 32)


def stage_flops(K, Np, Nfp, nvar):
    """Contraction flops of one stage, 2 K nvar Np (3 Np + 4 Nfp) (the model of
    the reference's calc_ref, utils/src/python/gpu_profile.py:8-31)."""
    return 2.0 * K * nvar * Np * (3 * Np + 4 * Nfp)


def cpu_baseline(pb, u0, seconds_target=15.0):
    """Time the oracle (C restatement, OpenMP) on a bounded number of steps."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as O

    try:
        ncpu = len(os.sched_getaffinity(0))
    except AttributeError:
        ncpu = os.cpu_count() or 1
    threads = max(1, min(16, ncpu))
    os.environ["OMP_NUM_THREADS"] = str(threads)
    o = O.Oracle(pb, "euler", cfl=0.4)
    t0 = time.perf_counter()
    o.run(u0, 1)
    t1 = time.perf_counter()
    per_step = t1 - t0
    n = max(1, min(10, int(seconds_target / max(per_step, 1e-3))))
    t0 = time.perf_counter()
    o.run(u0, n)
    wall = time.perf_counter() - t0
    dof = pb.K * pb.Np * 5
    return {"value": dof * 5 * n / wall, "unit": "DOF-stage updates/s", "cores": threads, "kind": "port",
            "sample": f"{n} LSRK4(5) steps of the same mesh and state (K={pb.K}, N={pb.N}, the reference LM flux) by oracle/dglm_oracle.c, "
                      f"OpenMP {threads} threads, after 1 warm-up step"}


def run_forward(pb, u0, precision="fp64", flux="rusanov", steps=200, warmup=10, kernel_steps=20,
                stage_kernel="auto"):
    """Warm up, time `steps` graph-replayed steps, then time `kernel_steps`
    eager steps with per-launch HIP events."""
    from exadglm_amd.solver import Solver

    s = Solver(pb, "euler", cfl=0.4, tmax=math.inf, precision=precision, flux=flux)
    if precision == "fp64":
        s.set_stage_kernel(stage_kernel)
    name = s.stage_kernel()
    s.set_state(u0)
    s.step(warmup)
    s.sync()
    t0 = time.perf_counter()
    s.step(steps)
    s.sync()
    wall = time.perf_counter() - t0
    t, dt, ts, _ = s.time()  # raises on a non-finite dt
    u = s.get_state()
    if not np.isfinite(u).all():
        raise RuntimeError("state is not finite after the timed steps")
    tm = None
    if kernel_steps:
        s.set_timing(True)
        s.step(kernel_steps)
        s.sync()
        tm = s.timing()
    s.close()
    return {"wall": wall, "u": u, "t": t, "kernel": name, "timing": tm}


def profiled_traffic(ksym, args):
    """HBM bytes per launch of the stage kernel from the committed rocprofv3
    summary of this workload (FETCH_SIZE x 2 + WRITE_SIZE x 1 KiB per dispatch,
    factors from a plain copy of known size), when the summary covers this
    kernel and configuration; (None, None) otherwise."""
    path = os.path.join("profiles", f"stage_ws_kernel_euler_N{args.order}_K{6 * args.nhex ** 3}.txt")
    if ksym != f"stage_ws_kernel<{args.order}, 0, false>" or args.precision != "fp64" or \
            not os.path.exists(os.path.join(ROOT, path)):
        return None, None
    in_table = False  # the memory-counter table: ... disp, FETCH KiB, WRITE KiB, HBM MB
    for line in open(os.path.join(ROOT, path)):
        if "FETCH KiB" in line and "WRITE KiB" in line:
            in_table = True
        elif in_table and line.startswith(f"void dglm::stage_ws_kernel<{args.order}, 0, false>"):
            f = line.split()
            try:
                fetch, write = float(f[-3]), float(f[-2])
            except ValueError:
                return None, None
            if fetch > 0 and write > 0:
                return (2 * fetch + write) * 1024, path + " (rocprofv3 --pmc FETCH_SIZE x2 + WRITE_SIZE x1 KiB)"
    return None, None


def per_var_err(u, ref):
    """Relative L2 per conserved variable; rho*w (zero in the exact z-extruded
    vortex) is measured against the norm of the whole momentum."""
    den = [np.linalg.norm(ref[1:4]) if v == 3 else np.linalg.norm(ref[v]) for v in range(5)]
    return [float(np.linalg.norm(u[v] - ref[v]) / den[v]) for v in range(5)]


def mixed_precision_leg(order=6, nhex=24, steps=490, warmup=10):
    from exadglm_amd.problem import isentropic_vortex, uniform_problem, vortex_box

    pb = uniform_problem(order, nhex + 1, nhex + 1, nhex + 1, box=vortex_box(nhex))
    a = pb.arrays
    u0 = isentropic_vortex(a["px"], a["py"], a["pz"])
    dof_stage = pb.K * pb.Np * 5 * 5 * steps
    r64 = run_forward(pb, u0, "fp64", "rusanov", steps, warmup, 0)
    r32 = run_forward(pb, u0, "fp32", "rusanov", steps, warmup, 0)
    exact = isentropic_vortex(a["px"], a["py"], a["pz"], t=r64["t"])
    split = {}
    for prec in ("bf16x3", "bf16x6"):
        try:
            rb = run_forward(pb, u0, prec, "rusanov", steps, warmup, 0)
        except Exception as e:  # noqa: BLE001  (reported in its entry; the FP64 and FP32 figures stand)
            print(f"bench: split-BF16 {prec} failed: {type(e).__name__}: {e}", file=sys.stderr, flush=True)
            split[prec] = {"error": f"{type(e).__name__}: {e}"}
            continue
        split[prec] = {"ms_per_step": rb["wall"] / steps * 1e3, "speedup_vs_fp64": r64["wall"] / rb["wall"],
                       "kernel": rb["kernel"], "rel_l2_vs_fp64": per_var_err(rb["u"], r64["u"])}
        del rb
    return {
        "workload": f"euler_isentropic_vortex_N{order}_{nhex}cube_hexes_x6_K{pb.K}_rusanov_{warmup + steps}steps",
        "fp64": {"value": dof_stage / r64["wall"], "ms_per_step": r64["wall"] / steps * 1e3, "kernel": r64["kernel"]},
        "mixed": {"value": dof_stage / r32["wall"], "ms_per_step": r32["wall"] / steps * 1e3, "kernel": r32["kernel"],
                  "contraction": "FP32 MFMA on the remainders of per-block least-squares affine fits of the FP64 "
                                 "fluxes, FP32 accumulation; the fits and their contraction (FP64 MFMA), state, "
                                 "pressure, wave speed and sums FP64"},
        "speedup": r64["wall"] / r32["wall"],
        "rel_l2_mixed_vs_fp64": per_var_err(r32["u"], r64["u"]),
        "rel_l2_fp64_vs_analytic": per_var_err(r64["u"], exact),
        "rel_l2_mixed_vs_analytic": per_var_err(r32["u"], exact),
        "split_bf16": dict(split, contraction="bf16 MFMA on the same affine-fit remainders cut into bf16 pieces: "
                                              "x3 = hi*hi + hi*lo + lo*hi, x6 = three pieces, six products; FP32 "
                                              "accumulation, generic per-variable kernel"),
        "unit": "DOF-stage updates/s",
        "note": "rho*w errors relative to the momentum norm; target rel L2 <= 1e-8 at >= 1.8x",
    }


def adjoint_leg(precision="fp64", nhex=32, order=4, steps=300, stage_kernel=None, forward_only=False):
    """Euler adjoint: J = int 1/2 rho |v|^2 at T; forward wall time (graph
    replay) against checkpointed recording + reverse sweep.  stage_kernel
    forces a stage kernel (the FP32-MFMA forward of the mixed-precision
    configuration: "mfma_mx"); forward_only times the forward alone."""
    from exadglm_amd.solver import Solver
    from exadglm_amd.workloads import c5, kinetic_energy, perturbed_state

    w = c5(nhex, order)
    pb = w["problem"]
    u0, _ = perturbed_state(w["base"], w["bump"])
    every = w["every"]
    s = Solver(pb, "euler", cfl=0.4, tmax=math.inf, precision=precision, flux="lm")
    if stage_kernel:
        s.set_stage_kernel(stage_kernel)
    kernel = s.stage_kernel()
    s.set_state(u0)
    s.step(2)
    s.set_state(u0)
    s.set_time(0.0, 0)
    s.sync()
    t0 = time.perf_counter()
    s.step(steps)
    s.sync()
    t_fwd = time.perf_counter() - t0
    if forward_only:
        s.close()
        return {"t_forward_s": t_fwd, "kernel": kernel, "K": pb.K}
    s.set_state(u0)
    s.set_time(0.0, 0)
    s.sync()
    t0 = time.perf_counter()
    s.adjoint_record(steps, every)
    s.sync()
    t_rec = time.perf_counter() - t0
    uT = s.get_state()
    tape = s.tape_bytes()  # the reverse sweep releases the kept segments
    J, g = kinetic_energy(pb, uT)
    t0 = time.perf_counter()
    grad = s.adjoint_backward(g)
    t_bwd = time.perf_counter() - t0
    s.close()
    return {"J": J, "grad": grad, "t_forward_s": t_fwd, "t_record_s": t_rec, "t_backward_s": t_bwd,
            "tape_bytes": tape, "K": pb.K, "every": every, "kernel": kernel}


def hw_check():
    """Device figures behind the roofline peaks, measured: compute units and
    LDS per workgroup from the runtime, HBM bandwidth from a 4 GiB device copy
    (read + write) and a 4 GiB read-only sum, and the rate of large library
    GEMMs (8192^3; a lower bound on the dense MFMA peak)."""
    import torch

    dev = torch.cuda.current_device()
    prop = torch.cuda.get_device_properties(dev)
    out = {"name": prop.name, "compute_units": prop.multi_processor_count,
           "lds_per_block_bytes": getattr(prop, "shared_memory_per_block", None),
           "hbm_total_bytes": prop.total_memory}

    def timed(fn, reps):
        fn()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            fn()
        e1.record()
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / 1e3 / reps

    n = (4 << 30) // 8
    src = torch.rand(n, dtype=torch.float64, device="cuda")
    dst = torch.empty_like(src)
    t = timed(lambda: dst.copy_(src), 10)
    out["hbm_copy_GBs"] = 2 * n * 8 / t / 1e9
    t = timed(lambda: src.sum(), 10)
    out["hbm_read_GBs"] = n * 8 / t / 1e9
    del src, dst
    m = 8192
    for name, dt in (("fp64", torch.float64), ("fp32", torch.float32), ("bf16", torch.bfloat16)):
        a = torch.rand(m, m, device="cuda").to(dt)
        b = torch.rand(m, m, device="cuda").to(dt)
        t = timed(lambda: torch.mm(a, b), 5)
        out[f"gemm_{name}_TFs"] = 2 * m ** 3 / t / 1e12
        del a, b
    torch.cuda.empty_cache()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--nhex", type=int, default=32)
    ap.add_argument("--order", type=int, default=4)
    ap.add_argument("--flux", default="rusanov", choices=["rusanov", "lm"],
                    help="headline flux (the workload's Rusanov flux; lm = the reference's)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-extra", action="store_true",
                    help="skip the LM-flux, mixed-precision and adjoint legs")
    ap.add_argument("--kernel-steps", type=int, default=20,
                    help="eager steps after the timed region with per-launch HIP event timing (roofline)")
    ap.add_argument("--stage-kernel", default="auto", choices=["auto", "simple", "mfma", "mfma_ws"])
    ap.add_argument("--precision", default="fp64", choices=["fp64", "fp32"],
                    help="contraction precision of the headline run")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB MiB three times (scales the FETCH_SIZE / "
                         "WRITE_SIZE counters of a profiled run to bytes)")
    ap.add_argument("--hw-check", action="store_true",
                    help="only measure the device figures the rooflines assume (CUs, LDS, HBM copy and read "
                         "bandwidth, library GEMM rates in FP64/FP32/BF16) and print them")
    ap.add_argument("--mixed-only", type=int, default=0, metavar="STEPS",
                    help="only run the FP64 and mixed-precision N=6 vortex legs with STEPS timed steps each "
                         "(for a profiler) and print their times")
    ap.add_argument("--adjoint-only", type=int, default=0, metavar="STEPS",
                    help="only run the FP64 Euler adjoint leg with STEPS steps (for a profiler) and print its times")
    args = ap.parse_args()
    if args.adjoint_only:
        res = adjoint_leg("fp64", steps=args.adjoint_only)
        res.pop("grad")
        res["ratio_record_plus_backward_over_forward"] = (res["t_record_s"] + res["t_backward_s"]) / \
            res["t_forward_s"]
        print(json.dumps(res), flush=True)
        return
    if args.mixed_only:
        res = mixed_precision_leg(steps=args.mixed_only, warmup=2)
        print(json.dumps({k: res[k] for k in ("workload", "fp64", "mixed", "speedup")}), flush=True)
        return
    if args.hw_check:
        print(json.dumps(hw_check()), flush=True)
        return
    if args.calibrate_copy:
        import torch

        n = args.calibrate_copy * (1 << 20) // 8
        src = torch.rand(n, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": n * 8, "copy_bytes_written": n * 8, "launches": 3}), flush=True)
        return
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (got --gpus {args.gpus}); the multi-GPU halo path is not built",
              file=sys.stderr)
        sys.exit(2)

    from exadglm_amd.problem import isentropic_vortex, uniform_problem, vortex_box

    n1 = args.nhex + 1
    pb = uniform_problem(args.order, n1, n1, n1, box=vortex_box(args.nhex))
    a = pb.arrays
    u0 = isentropic_vortex(a["px"], a["py"], a["pz"])
    K, Np, Nfp, nvar = pb.K, pb.Np, pb.Nfp, 5

    r = run_forward(pb, u0, args.precision, args.flux, args.steps, args.warmup, args.kernel_steps,
                    args.stage_kernel)
    wall, tm, kernel_name = r["wall"], r["timing"], r["kernel"]
    dof_stage = K * Np * nvar * 5 * args.steps
    value = dof_stage / wall
    # Stage-kernel time inside the timed region: the graph-replayed step time
    # less the per-step wave-speed and dt kernels (their event-timed averages
    # from the eager steps after it), over the 5 stages.  The eager steps'
    # own per-launch stage average is reported beside it.
    st_ms, st_n = tm["stage"]
    st_eager_s = st_ms / st_n / 1e3
    other_s = sum(tm[k][0] / tm[k][1] / 1e3 for k in ("wavespeed", "dt") if tm[k][1])
    st_avg_s = (wall / args.steps - other_s) / 5
    B = stage_bytes(K, Np, Nfp, nvar)
    F = stage_flops(K, Np, Nfp, nvar)
    gbs = B / st_avg_s / 1e9
    tfl = F / st_avg_s / 1e12
    pk = FP32_PEAK_TFLOPS if args.precision == "fp32" else FP64_PEAK_TFLOPS
    t_hbm = B / (HBM_PEAK_GBS * 1e9)
    t_fp = F / (pk * 1e12)
    if t_hbm >= t_fp:
        roof = {"bound": "hbm", "achieved": gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": gbs / HBM_PEAK_GBS}
    else:
        roof = {"bound": "mfma", "achieved": tfl, "peak": pk, "unit": "TFLOP/s", "frac": tfl / pk}
    ksym = f"stage_ws_kernel<{args.order}, 0, false>" if kernel_name == "mfma_ws" else \
        f"stage_{kernel_name}_kernel<{args.order}, EULER, 0>"
    roof["traffic"], roof["traffic_source"] = profiled_traffic(ksym, args)
    roof["kernel"] = ksym
    roof["avg_launch_us"] = st_avg_s * 1e6
    roof["avg_launch_us_source"] = ("timed graph-replayed steps: (ms_per_step - wave-speed and dt kernel "
                                    "averages) / 5 stages")
    roof["avg_launch_us_eager_events"] = st_eager_s * 1e6
    roof["algorithmic_bytes_per_launch"] = B
    roof["contraction_flops_per_launch"] = F
    roof["other"] = ({"unit": "TFLOP/s", "achieved": tfl, "peak": pk, "frac": tfl / pk}
                     if roof["bound"] == "hbm" else
                     {"unit": "GB/s", "achieved": gbs, "peak": HBM_PEAK_GBS, "frac": gbs / HBM_PEAK_GBS})
    roof["binding_roof_us"] = max(t_hbm, t_fp) * 1e6
    roof["frac_of_binding_roof"] = max(t_hbm, t_fp) / st_avg_s

    out = {
        "metric": "DOF-stage updates/s: FP64 vs mixed @L2<=1e-8; adjoint/forward time",
        "value": value,
        "unit": "DOF-stage updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else "f32 contraction, f64 state",
        "data": "synthetic",
        "config": {"workload": f"euler_isentropic_vortex_N{args.order}_{args.nhex}cube_hexes_x6_K{K}_"
                               f"{args.flux}_{args.precision}",
                   "K": K, "N": args.order, "Np": Np, "nvar": nvar, "cfl": 0.4, "precision": args.precision,
                   "flux": args.flux, "stepper": "LSRK4(5)"},
        "roofline": roof,
        "kernel_ms": {k: {"total_ms": v[0], "launches": v[1]} for k, v in tm.items()},
    }
    if not args.no_extra:
        # the secondary figures: a failure in one of them is reported in its
        # entry (and on stderr) instead of losing the whole line
        def leg(name, fn):
            try:
                out[name] = fn()
            except Exception as e:  # noqa: BLE001
                print(f"bench: {name} failed: {type(e).__name__}: {e}", file=sys.stderr, flush=True)
                out[name] = {"error": f"{type(e).__name__}: {e}"}

        other = "lm" if args.flux == "rusanov" else "rusanov"
        del r

        def flux_leg():
            r2 = run_forward(pb, u0, args.precision, other, args.steps, args.warmup, args.kernel_steps,
                             args.stage_kernel)
            st2 = r2["timing"]["stage"]
            return {"value": dof_stage / r2["wall"], "ms_per_step": r2["wall"] / args.steps * 1e3,
                    "stage_avg_us": st2[0] / st2[1] * 1e3, "kernel": r2["kernel"]}

        def adjoint_legs():
            adj = {}
            grads = {}
            # the mixed-precision configuration's FP32-MFMA forward is the fused
            # mixed kernel, selected explicitly: an FP32 request at N = 4 (AUTO)
            # runs the FP64 warp-specialised kernel, which is faster there
            for prec in ("fp64", "fp32"):
                res = adjoint_leg(prec, stage_kernel="mfma_mx" if prec == "fp32" else None)
                grads[prec] = res.pop("grad")
                res["ratio_backward_over_forward"] = res["t_backward_s"] / res["t_forward_s"]
                res["ratio_record_plus_backward_over_forward"] = (res["t_record_s"] + res["t_backward_s"]) / \
                    res["t_forward_s"]
                adj[f"forward_{prec}"] = res
            adj["grad_rel_diff_fp32_forward_vs_fp64"] = float(np.linalg.norm(grads["fp32"] - grads["fp64"]) /
                                                              np.linalg.norm(grads["fp64"]))
            adj["forward_fp32_auto"] = adjoint_leg("fp32", forward_only=True)
            adj["workload"] = "euler_adjoint_vortex_plus_1e-3_gaussian_bumps_N4_32cube_K196608_300steps_" \
                              "J=int(rho|v|^2/2)(T), adjoint FP64"
            return adj

        leg(f"{other}_flux", flux_leg)
        leg("mixed_precision", mixed_precision_leg)
        leg("adjoint", adjoint_legs)
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(pb, u0)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
