"""The discrete adjoint on the GPU: dot-product tests against the oracle's
tangent-linear RHS, trajectory dot products (advection), central finite
differences of the real forward solver (Euler, including the tau(u) and
dt(u) paths), checkpoint invariance, and the torch.autograd binding."""
import math

import numpy as np
import pytest

from helpers import rel_err, smooth_state
from exadglm_amd.problem import cube_problem, isentropic_vortex, uniform_problem, vortex_box
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
    """taubar = (dR/dtau)^T rbar against differences of the oracle RHS in tau."""
    _say("tau adjoint")
    pb = cube_problem(3, 3)
    u = smooth_state(pb, "euler", seed=61)
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


def _euler_setup(N=3, nh=4):
    pb = uniform_problem(N, nh + 1, nh + 1, nh + 1, box=vortex_box(nh))
    a = pb.arrays
    u0 = isentropic_vortex(a["px"], a["py"], a["pz"])
    return pb, u0


def _objective(u, wts):
    return 0.5 * float(np.sum(wts * u * u))


def test_euler_gradient_vs_finite_differences(gpu):
    """J = 1/2 sum w u(T)^2 after 20 steps: adjoint gradient vs central
    differences (h = 1e-6 relative) of the actual forward solver along random
    directions; the exact adjoint follows the tau and dt argmax paths."""
    _say("Euler gradient vs FD")
    pb, u0 = _euler_setup()
    rng = np.random.default_rng(5)
    wts = rng.uniform(0.5, 1.5, u0.shape)
    nsteps = 20
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

        fails = 0
        for q in range(4):
            d = rng.standard_normal(u0.shape) * np.abs(u0)
            h = 1e-6
            fd = (J(u0 + h * d) - J(u0 - h * d)) / (2 * h)
            ad = dot(grad, d)
            err = abs(fd - ad) / abs(fd)
            _say(f"  direction {q}: fd={fd:.12e} adjoint={ad:.12e} rel={err:.2e}")
            fails += err > 1e-6
    assert fails == 0


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
    """Freezing tau and dt changes the Euler gradient (they depend on u) but
    not the advection one (they do not)."""
    _say("frozen-coefficient adjoint")
    pb, u0 = _euler_setup()
    w = np.random.default_rng(2).standard_normal(u0.shape)
    out = []
    for exact in (True, False):
        with Solver(pb, "euler", cfl=0.4) as s:
            s.set_adjoint_exact(exact)
            s.set_state(u0)
            s.adjoint_record(6, 3)
            out.append(s.adjoint_backward(w))
    assert rel_err(out[1], out[0]) > 1e-8


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
