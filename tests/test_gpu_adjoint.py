"""The discrete adjoint on the GPU: dot-product tests against the oracle's
tangent-linear RHS, trajectory dot products (advection), central finite
differences of the real forward solver (Euler, including the tau(u) and
dt(u) paths), checkpoint invariance, and the torch.autograd binding."""

import numpy as np
import pytest

from helpers import noisy_state, rel_err, smooth_state
from traj_tl import Trajectory
from exadglm_amd.problem import cube_problem
from exadglm_amd.solver import Solver
import pyoracle as O

pytestmark = pytest.mark.gpu
PHYS = ["advection", "euler"]


def _say(msg):
    print(msg, flush=True)


def dot(a, b):
    return float(np.dot(np.ravel(a), np.ravel(b)))


@pytest.mark.parametrize("physics", PHYS)
@pytest.mark.parametrize("N", [1, 2, 3, 4, 5, 6])
def test_adjoint_rhs_dot_product(gpu, physics, N):
    """<J du, rbar> (oracle tangent-linear, tau frozen) = <du, J^T rbar> (GPU)."""
    _say(f"adjoint rhs dot {physics} N={N}")
    pb = cube_problem(N, 3)
    u = smooth_state(pb, physics, seed=N + 51)
    rng = np.random.default_rng(N)
    du = rng.standard_normal(u.shape)
    rbar = rng.standard_normal(u.shape)
    with Solver(pb, physics) as s:
        ubar, _ = s.adjoint_rhs(u, rbar)
    tl = O.rhs_tl(O.Oracle(pb, physics), u, du)
    lhs, rhs = dot(tl, rbar), dot(du, ubar)
    scale = np.linalg.norm(tl) * np.linalg.norm(rbar)
    assert abs(lhs - rhs) <= 1e-13 * scale, (lhs, rhs)


def test_adjoint_tau_sensitivity(gpu):
    """taubar = (dR/dtau)^T rbar against differences of the oracle RHS in tau,
    on a discontinuous state (on a continuous one u- = u+ and the tau term
    vanishes)."""
    _say("tau adjoint")
    pb = cube_problem(3, 3)
    u = noisy_state(pb, "euler", seed=61)
    rng = np.random.default_rng(7)
    rbar = rng.standard_normal(u.shape)
    dtau = rng.standard_normal((pb.K, 4))
    with Solver(pb, "euler") as s:
        _, taubar = s.adjoint_rhs(u, rbar)
    o = O.Oracle(pb, "euler")
    _, _, tau = o.rhs(u)
    # R is affine in tau: the difference is exact up to rounding
    d = O.rhs_tau(o, u, tau + dtau) - O.rhs_tau(o, u, tau)
    lhs, rhs = dot(d, rbar), dot(taubar, dtau)
    assert abs(lhs - rhs) <= 1e-11 * np.linalg.norm(d) * np.linalg.norm(rbar), (lhs, rhs)


def test_advection_trajectory_dot_product(gpu):
    """Linear advection: <S du, w> = <du, S^T w> over 40 LSRK steps."""
    _say("advection trajectory dot product")
    pb = cube_problem(3, 4)
    rng = np.random.default_rng(11)
    n = pb.K * pb.Np
    du = rng.standard_normal((1, n))
    w = rng.standard_normal((1, n))
    with Solver(pb, "advection", cfl=0.5) as s:
        s.set_velocity(1.0, 0.5, 0.25)
        s.set_state(du)
        s.adjoint_record(40, 7)
        Sdu = s.get_state()
        STw = s.adjoint_backward(w)
    lhs, rhs = dot(Sdu, w), dot(du, STw)
    assert abs(lhs - rhs) <= 1e-12 * np.linalg.norm(Sdu) * np.linalg.norm(w), (lhs, rhs)


def _euler_setup(N=3, nh=3):
    """A small Euler problem on a state without wave-speed ties, where the
    trajectory is differentiable (tau and dt maxima at single nodes)."""
    pb = cube_problem(N, nh)
    return pb, noisy_state(pb, "euler", seed=3)


def _objective(u, wts):
    return 0.5 * float(np.sum(wts * u * u))


def test_euler_gradient_vs_tangent_linear_and_fd(gpu):
    """J = 1/2 sum w u(T)^2 after 20 steps.  The adjoint gradient along random
    directions equals the CPU tangent-linear of the oracle trajectory (exact
    derivative of the same discrete map, tau(u) and dt(u) paths included) to
    rounding, and central differences (h = 1e-6 relative) of the GPU forward
    itself to truncation."""
    _say("Euler gradient vs tangent-linear and FD")
    pb, u0 = _euler_setup()
    rng = np.random.default_rng(5)
    wts = rng.uniform(0.5, 1.5, u0.shape)
    nsteps = 20
    T = Trajectory(pb, "euler", cfl=0.4)
    with Solver(pb, "euler", cfl=0.4) as s:
        s.set_state(u0)
        s.adjoint_record(nsteps, 5)
        uT = s.get_state()
        grad = s.adjoint_backward(wts * uT)

        def J(x):
            s.set_state(x)
            s.set_time(0.0, 0)
            s.step(nsteps)
            return _objective(s.get_state(), wts)

        for q in range(3):
            d = rng.standard_normal(u0.shape) * np.abs(u0)
            uTc, duT = T.run(u0, nsteps, d)
            tl = dot(wts * uTc, duT)
            h = 1e-6
            fd = (J(u0 + h * d) - J(u0 - h * d)) / (2 * h)
            ad = dot(grad, d)
            _say(f"  direction {q}: tl={tl:.15e} adjoint={ad:.15e} fd={fd:.12e}")
            assert abs(ad - tl) <= 1e-10 * abs(tl), (ad, tl)
            assert abs(ad - fd) <= 1e-6 * abs(fd), (ad, fd)


def test_checkpoint_interval_does_not_change_gradient(gpu):
    _say("checkpoint invariance")
    pb, u0 = _euler_setup()
    wts = np.random.default_rng(9).uniform(0.5, 1.5, u0.shape)
    grads = []
    for every in (1, 4, 12):
        with Solver(pb, "euler", cfl=0.4) as s:
            s.set_state(u0)
            s.adjoint_record(12, every)
            uT = s.get_state()
            grads.append(s.adjoint_backward(wts * uT))
    assert np.array_equal(grads[0], grads[1]) and np.array_equal(grads[0], grads[2])


def test_frozen_coefficient_mode(gpu):
    """With tau and dt frozen the adjoint is the transpose of the CPU
    tangent-linear with frozen coefficients (dot product to rounding), and it
    differs from the exact-mode gradient."""
    _say("frozen-coefficient adjoint")
    pb, u0 = _euler_setup()
    rng = np.random.default_rng(2)
    w = rng.standard_normal(u0.shape)
    d = rng.standard_normal(u0.shape) * np.abs(u0)
    out = []
    for exact in (True, False):
        with Solver(pb, "euler", cfl=0.4) as s:
            s.set_adjoint_exact(exact)
            s.set_state(u0)
            s.adjoint_record(6, 3)
            out.append(s.adjoint_backward(w))
    assert rel_err(out[1], out[0]) > 1e-8
    T = Trajectory(pb, "euler", cfl=0.4)
    for g, frozen in zip(out, (False, True)):
        _, duT = T.run(u0, 6, d, frozen=frozen)
        tl = dot(w, duT)
        assert abs(dot(g, d) - tl) <= 1e-10 * abs(tl), (frozen, dot(g, d), tl)


def test_torch_autograd(gpu):
    import torch

    from exadglm_amd.autograd import dglm_evolve

    _say("torch autograd")
    pb, u0 = _euler_setup()
    wts = np.random.default_rng(5).uniform(0.5, 1.5, u0.shape)
    with Solver(pb, "euler", cfl=0.4) as s:
        x = torch.tensor(u0, dtype=torch.float64, device="cuda", requires_grad=True)
        w = torch.tensor(wts, dtype=torch.float64, device="cuda")
        uT = dglm_evolve(x, s, 10, every=3)
        J = 0.5 * (w * uT * uT).sum()
        J.backward()
        g = x.grad.cpu().numpy()
        s.set_state(u0)
        s.set_time(0.0, 0)
        s.adjoint_record(10, 3)
        g2 = s.adjoint_backward(wts * s.get_state())
    assert np.array_equal(g, g2)


def test_c4_advection_adjoint_full_size(gpu):
    """The advection adjoint configuration at full size: periodic cube of 20^3
    hexes x 6 (K = 48,000), N = 5, v = (1, 0.5, 0.25), random Fourier field
    u0 (seed 0), target = the forward run of the seed-1 field, J = 1/2
    ||u(T) - u_target||^2 (element mass matrices) after 400 LSRK steps,
    checkpoints every 20, all FP64.  Checks: the dot-product test of the
    400-step map, central differences (h = 1e-6) along 8 random directions
    (J is quadratic, so they are exact up to rounding), and the
    backward-transport gradient: for constant-velocity transport on the
    periodic box S^T M S ~ M, so M^-1 dJ/du0 ~ u0 - u1 (seed-0 minus seed-1
    field) up to the discretisation error."""
    from exadglm_amd.problem import fourier_modes
    from exadglm_amd.workloads import c4, l2_misfit

    def mode_field(seed):
        """Random combination of the 16 Fourier modes of u0 and u1 (fields
        built from other modes are orthogonal to both, and every inner product
        of the test would vanish)."""
        k = np.concatenate([fourier_modes(0)[0], fourier_modes(1)[0]])
        rng = np.random.default_rng(seed)
        x, y, z = a["px"].ravel(), a["py"].ravel(), a["pz"].ravel()
        u = np.zeros_like(x)
        for km in k:
            u += rng.standard_normal() * np.cos(2 * np.pi * (km[0] * x + km[1] * y + km[2] * z)
                                                + rng.uniform(0, 2 * np.pi))
        return u[None, :]

    w = c4()
    pb = w["problem"]
    assert pb.K == 48000 and pb.Np == 56
    a = pb.arrays
    steps, every = w["steps"], w["every"]
    _say(f"C4 advection adjoint K={pb.K} N=5 {steps} steps, checkpoints every {every}")
    u0, u1 = w["u0"], w["u_target0"]
    with Solver(pb, "advection", cfl=w["cfl"]) as s:
        s.set_velocity(*w["velocity"])

        def fwd(u):
            s.set_state(u)
            s.set_time(0.0, 0)
            s.step(steps)
            return s.get_state()

        ut = fwd(u1)
        s.set_state(u0)
        s.set_time(0.0, 0)
        s.adjoint_record(steps, every)
        uT = s.get_state()
        J0, gT = l2_misfit(pb, uT, ut)
        grad = s.adjoint_backward(gT)
        # dot-product test of the 400-step map with two smooth random fields
        du, wv = mode_field(20), mode_field(21)
        s.set_state(du)
        s.set_time(0.0, 0)
        s.adjoint_record(steps, every)
        Sdu = s.get_state()
        STw = s.adjoint_backward(wv)
        lhs, rhs = dot(Sdu, wv), dot(du, STw)
        _say(f"  dot product {lhs:.15e} {rhs:.15e}")
        assert abs(lhs - rhs) <= 1e-12 * np.linalg.norm(Sdu) * np.linalg.norm(wv)
        # central differences along 8 directions
        h = 1e-6
        for q in range(8):
            d = mode_field(30 + q)
            fd = (l2_misfit(pb, fwd(u0 + h * d), ut)[0] - l2_misfit(pb, fwd(u0 - h * d), ut)[0]) / (2 * h)
            ad = dot(grad, d)
            _say(f"  direction {q}: adjoint={ad:.12e} fd={fd:.12e}")
            assert abs(ad - fd) <= 1e-7 * abs(fd), (ad, fd)
    # backward transport: M^-1 grad against u0 - u1
    Minv = np.linalg.inv(pb.op_ref.M)
    g = (grad.reshape(pb.K, pb.Np) @ Minv.T) / pb.mesh.J[:, None]
    diff = (u0 - u1).reshape(pb.K, pb.Np)
    e = rel_err(g, diff)
    _say(f"  J={J0:.6e}  |M^-1 grad - (u0 - u1)| / |u0 - u1| = {e:.3e}")
    assert e < 1e-5  # measured 2.1e-6: the discrete adjoint is backward transport to that level


def test_c5_euler_adjoint_full_size(gpu):
    """The Euler adjoint configuration at full size: the C2 vortex (K = 196,608,
    N = 4) with a 1e-3 Gaussian-bump field eps added to rho and p, J = int
    1/2 rho |v|^2 at T after 300 steps (LM flux), sqrt(300) ~ 17 checkpoints,
    gradient with respect to eps.  Through torch.autograd (eps -> u0 ->
    dglm_evolve -> J): FP64 forward against central differences of the FP64
    forward along 3 random bump fields, and the FP32-MFMA forward with the FP64
    adjoint against the all-FP64 gradient."""
    import torch

    from exadglm_amd.autograd import dglm_evolve
    from exadglm_amd.problem import gaussian_bumps, vortex_box
    from exadglm_amd.workloads import GAMMA, c5, kinetic_energy, mass_weights, perturbed_state

    w = c5()
    pb = w["problem"]
    a = pb.arrays
    assert pb.K == 196608
    steps, every = w["steps"], w["every"]
    base = w["base"]
    dev = "cuda"
    rho, mx, my, mz, E = (torch.tensor(x, device=dev) for x in base)
    q2 = (mx * mx + my * my + mz * mz) / (rho * rho)
    p = (GAMMA - 1) * (E - 0.5 * rho * q2)
    wq = torch.tensor(mass_weights(pb), device=dev)

    def objective(solver, eps):
        r2, p2 = rho + eps, p + eps
        u0 = torch.stack([r2, r2 * mx / rho, r2 * my / rho, r2 * mz / rho, p2 / (GAMMA - 1) + 0.5 * r2 * q2])
        uT = dglm_evolve(u0, solver, steps, every)
        return (wq * 0.5 * (uT[1] ** 2 + uT[2] ** 2 + uT[3] ** 2) / uT[0]).sum()

    grads = {}
    for prec in ("fp64", "fp32"):
        _say(f"C5 Euler adjoint K={pb.K} N=4 {steps} steps, {prec} forward")
        with Solver(pb, "euler", cfl=w["cfl"], precision=prec) as s:
            if prec == "fp32":
                # the FP32-MFMA forward is the fused mixed kernel, selected
                # explicitly (AUTO runs the faster FP64 kernel at N = 4)
                assert s.stage_kernel() == "mfma_ws"
                s.set_stage_kernel("mfma_mx")
            eps = torch.tensor(w["bump"], device=dev, requires_grad=True)
            J = objective(s, eps)
            J.backward()
            grads[prec] = eps.grad.cpu().numpy()
            if prec == "fp64":
                # the C-ABI gradient and the numpy objective agree with the torch chain
                u0, dud = perturbed_state(base, w["bump"])
                s.set_state(u0)
                s.set_time(0.0, 0)
                s.adjoint_record(steps, every)
                Jn, g = kinetic_energy(pb, s.get_state())
                gu = s.adjoint_backward(g)
                assert abs(Jn - float(J)) <= 1e-12 * abs(Jn)
                assert rel_err((gu * dud).sum(axis=0), grads[prec]) < 1e-12
                with torch.no_grad():
                    for q in range(3):
                        d = gaussian_bumps(a["px"], a["py"], a["pz"], 10 + q, vortex_box(32))
                        h = 1e-3  # the bump amplitude is 1e-3: eps moves by ~1e-6
                        dt = torch.tensor(d, device=dev)
                        fd = (float(objective(s, eps + h * dt)) - float(objective(s, eps - h * dt))) / (2 * h)
                        ad = float(np.dot(grads[prec], d))
                        _say(f"  direction {q}: adjoint={ad:.12e} fd={fd:.12e}")
                        assert abs(ad - fd) <= 1e-6 * abs(fd), (ad, fd)
    e = rel_err(grads["fp32"], grads["fp64"])
    _say(f"  FP32-forward gradient vs FP64: {e:.3e}")
    assert 0 < e < 1e-5  # a different forward, close to the FP64 one


@pytest.mark.parametrize("nsteps", [20, 21])
def test_kept_segments_do_not_change_gradient(gpu, nsteps):
    """Segments kept whole by the recording (no recompute in the reverse sweep)
    give the gradient bitwise: none kept, two kept, and the default (as many
    as half the free memory holds, here all of them), also when the last
    segment is partial (21 steps, checkpoints every 4); the tape grows with
    them, the state after the recording is the same, and the reverse sweep
    releases the kept segments' memory (a second sweep of the same recording
    recomputes them and gives the same gradient)."""
    _say(f"kept segments, {nsteps} steps")
    pb, u0 = _euler_setup()
    wts = np.random.default_rng(7).uniform(0.5, 1.5, u0.shape)
    res = {}
    with Solver(pb, "euler", cfl=0.4) as s:
        for keep in (0, 2, -1):
            s.set_adjoint_keep(keep)
            s.set_state(u0)
            s.set_time(0.0, 0)
            s.adjoint_record(nsteps, 4)
            uT = s.get_state()
            tape = s.tape_bytes()
            g = s.adjoint_backward(wts * uT)
            res[keep] = (uT, g, tape, s.tape_bytes(), s.adjoint_backward(wts * uT))
    for keep in (2, -1):
        np.testing.assert_array_equal(res[keep][0], res[0][0])
        np.testing.assert_array_equal(res[keep][1], res[0][1])
        np.testing.assert_array_equal(res[keep][4], res[0][1])
        assert res[keep][3] == res[0][2]  # kept segments released by the sweep
    assert res[0][2] < res[2][2] < res[-1][2]


def test_failed_recording_leaves_no_tape(gpu):
    """An allocation failure while the recording sizes a new tape shape leaves
    an empty tape: the reverse sweep refuses it (DGLM_ERR_STATE) instead of
    reading freed buffers, and a retry records and differentiates normally.
    A failure to allocate the kept segments falls back to keeping none, with
    the gradient bitwise unchanged."""
    from exadglm_amd import _capi as A

    _say("failed recording")
    pb, u0 = _euler_setup()
    wts = np.random.default_rng(7).uniform(0.5, 1.5, u0.shape)
    with Solver(pb, "euler", cfl=0.4) as s:
        s.set_adjoint_keep(0)
        s.set_state(u0)
        s.set_time(0.0, 0)
        s.adjoint_record(12, 4)
        g_ref = s.adjoint_backward(wts * s.get_state())
        # a new shape whose first buffer cannot be allocated
        s.set_state(u0)
        s.set_time(0.0, 0)
        A.check(A.lib().dglm_set_alloc_faults(1))
        try:
            with pytest.raises(A.DGLMError) as ei:
                s.adjoint_record(12, 3)
            assert ei.value.code == A.DGLM_ERR_HIP
        finally:
            A.check(A.lib().dglm_set_alloc_faults(0))
        with pytest.raises(A.DGLMError) as ei:
            s.adjoint_backward(wts * u0)
        assert ei.value.code == A.DGLM_ERR_STATE
        # retry with the old shape: reallocated and recorded
        s.set_state(u0)
        s.set_time(0.0, 0)
        s.adjoint_record(12, 4)
        np.testing.assert_array_equal(s.adjoint_backward(wts * s.get_state()), g_ref)
        # kept segments that cannot be allocated: none kept, same gradient
        s.set_adjoint_keep(-1)
        s.set_state(u0)
        s.set_time(0.0, 0)
        A.check(A.lib().dglm_set_alloc_faults(1))
        try:
            s.adjoint_record(12, 4)
        finally:
            A.check(A.lib().dglm_set_alloc_faults(0))
        tape_none = s.tape_bytes()
        np.testing.assert_array_equal(s.adjoint_backward(wts * s.get_state()), g_ref)
        s.set_state(u0)
        s.set_time(0.0, 0)
        s.adjoint_record(12, 4)
        assert s.tape_bytes() > tape_none  # with the memory, segments are kept
        np.testing.assert_array_equal(s.adjoint_backward(wts * s.get_state()), g_ref)


def test_autograd_refuses_a_replaced_recording(gpu):
    """Two dglm_evolve calls on one solver before backward: the solver's tape
    holds the second trajectory, so backward through the first raises rather
    than returning the second trajectory's gradient; with a solver per
    evolve the chained gradient goes through."""
    import torch

    from exadglm_amd.autograd import dglm_evolve

    _say("autograd re-record guard")
    pb, u0 = _euler_setup()
    dev = torch.device("cuda:0")
    x = torch.tensor(u0, device=dev, requires_grad=True)
    with Solver(pb, "euler", cfl=0.4) as s:
        y = dglm_evolve(dglm_evolve(x, s, 4), s, 4)
        with pytest.raises(RuntimeError, match="re-recorded"):
            (y * y).sum().backward()
    x.grad = None
    with Solver(pb, "euler", cfl=0.4) as s1, Solver(pb, "euler", cfl=0.4) as s2:
        y = dglm_evolve(dglm_evolve(x, s1, 4), s2, 4, t0=0.0)
        (y * y).sum().backward()
    assert x.grad is not None and torch.isfinite(x.grad).all()
