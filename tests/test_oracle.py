"""Pins of the CPU oracle: against the compiled reference solver (oracle/_ref),
against the reference's published end-to-end L2 errors, and against
mathematical invariants (free stream, conservation, exact transport)."""
import glob
import math
import os

import numpy as np
import pytest

from helpers import noisy_state, rel_err, smooth_state
from exadglm_amd.problem import cube_problem, ic_advection_sine, ic_euler_density_wave, l2_error
import pyoracle as O

PHYS = ["advection", "euler"]


def _need_ref(physics, N):
    if not O.ref_available(physics, N):
        pytest.skip("oracle/_ref not built (needs the reference checkout)")


@pytest.mark.parametrize("physics", PHYS)
@pytest.mark.parametrize("N", list(range(1, 11)))
def test_rhs_matches_compiled_reference(physics, N):
    _need_ref(physics, N)
    pb = cube_problem(N, 2)
    u = smooth_state(pb, physics, seed=N)
    o = O.Oracle(pb, physics)
    r = O.RefSolver(pb, physics)
    rhs_o, mv_o, tau_o = o.rhs(u)
    rhs_r, mv_r, tau_r = r.rhs(u)
    assert rel_err(mv_o, mv_r) < 1e-15
    assert rel_err(tau_o, tau_r) < 1e-15
    assert rel_err(rhs_o, rhs_r) < 1e-13


@pytest.mark.parametrize("physics,N,nsteps", [("advection", 3, 20), ("euler", 4, 10)])
def test_steps_match_compiled_reference(physics, N, nsteps):
    _need_ref(physics, N)
    pb = cube_problem(N, 3)
    u0 = smooth_state(pb, physics, seed=3)
    uo, dto, to = O.Oracle(pb, physics, cfl=0.5).run(u0, nsteps)
    ur, dtr, tr = O.RefSolver(pb, physics, cfl=0.5).steps(u0, nsteps)
    assert rel_err(dto, dtr) < 1e-14
    for v in range(uo.shape[0]):
        assert rel_err(uo[v], ur[v]) < 1e-13


def _advection_exact(pb, t):
    m = pb.mesh
    x, y, z = m.high_PX.ravel(), m.high_PY.ravel(), m.high_PZ.ravel()
    xv, yv, zv = ((x - t) + 1) % 1 - 1, ((y - t) + 1) % 1 - 1, ((z - t) + 1) % 1 - 1
    return np.sin(2 * np.pi * xv) * np.sin(2 * np.pi * yv) * np.sin(2 * np.pi * zv)


def test_published_advection_l2():
    """ex1.advection.11.cpu.single.ipynb:214 -- K=1296, N=5, CFL 0.8, tmax 2:
    809 steps, L2 = 2.284502600150407e-05.  The reference's own CPU and GPU
    runs agree to ~2e-12 relative; machine-level FMA / ulp differences of the
    tables give ~1e-11 here, so the bound is 1e-9 relative."""
    pb = cube_problem(5, 6)
    u0 = ic_advection_sine(pb.arrays["px"], pb.arrays["py"], pb.arrays["pz"])
    u, dts, t = O.Oracle(pb, "advection", cfl=0.8, tmax=2.0).run(u0, 900)
    assert int((dts > 0).sum()) == 809 and t == 2.0
    l2 = l2_error(pb, _advection_exact(pb, 2.0), u[0])
    assert abs(l2 - 2.284502600150407e-05) / 2.284502600150407e-05 < 1e-9


@pytest.mark.slow
def test_published_euler_l2():
    """ex3.euler.11.cpu.single.ipynb:221 -- density wave, K=1296, N=5: 1318
    steps, L2(rho) = 2.429869462976010e-05 (bound as above)."""
    pb = cube_problem(5, 6)
    u0 = ic_euler_density_wave(pb.arrays["px"], pb.arrays["py"], pb.arrays["pz"])
    u, dts, t = O.Oracle(pb, "euler", cfl=0.8, tmax=2.0).run(u0, 1400)
    assert int((dts > 0).sum()) == 1318
    m = pb.mesh
    x, y, z = m.high_PX.ravel(), m.high_PY.ravel(), m.high_PZ.ravel()
    xv, yv, zv = ((x - 2) + 1) % 1 - 1, ((y + 1) + 1) % 1 - 1, ((z - 2) + 1) % 1 - 1
    ref = 1 + 0.2 * np.sin(2 * np.pi * (xv + yv + zv))
    l2 = l2_error(pb, ref, u[0])
    assert abs(l2 - 2.429869462976010e-05) / 2.429869462976010e-05 < 1e-9


@pytest.mark.parametrize("N", [2, 4, 6])
def test_free_stream_preserved(N):
    pb = cube_problem(N, 2)
    n = pb.K * pb.Np
    o = O.Oracle(pb, "euler")
    rho, vx, vy, vz, p = 1.3, 0.4, -0.2, 0.7, 2.0
    u = np.stack([np.full(n, c) for c in (rho, rho * vx, rho * vy, rho * vz,
                                          p / 0.4 + 0.5 * rho * (vx * vx + vy * vy + vz * vz))])
    rhs, _, _ = o.rhs(u)
    # round-off scale: |operator row| x |metric| x |flux|
    a = pb.arrays
    opn = max(np.abs(a[k]).sum(axis=1).max() for k in ("WDr", "WDs", "WDt", "LIFT"))
    met = max(np.abs(a[k]).max() for k in ("rx", "ry", "rz", "sx", "sy", "sz", "tx", "ty", "tz", "Fscale"))
    scale = 3 * opn * met * np.abs(u).max()
    assert np.abs(rhs).max() < 1e-14 * scale
    oa = O.Oracle(pb, "advection")
    rhs, _, _ = oa.rhs(np.full((1, n), 2.5))
    assert np.abs(rhs).max() < 1e-14 * 3 * opn * met * 2.5


@pytest.mark.parametrize("physics", PHYS)
def test_conservation_of_rhs(physics):
    """sum_e J_e 1^T M rhs_e = 0 on a periodic mesh (flux telescoping)."""
    pb = cube_problem(3, 3)
    u = smooth_state(pb, physics, seed=7)
    rhs, _, _ = O.Oracle(pb, physics).rhs(u)
    w = pb.op_ref.M.sum(axis=0)
    for v in range(rhs.shape[0]):
        tot = (pb.mesh.J[:, None] * (rhs[v].reshape(pb.K, pb.Np) * w)).sum()
        scale = (pb.mesh.J[:, None] * np.abs(rhs[v].reshape(pb.K, pb.Np)) * np.abs(w)).sum()
        assert abs(tot) < 1e-13 * scale


def test_planted_connectivity_error_is_caught():
    """A wrong node pairing across faces (vmapP rolled within each face) in the
    oracle's inputs breaks conservation: the invariant above is sensitive to
    that class of mistake."""
    pb = cube_problem(3, 3)
    u = smooth_state(pb, "euler", seed=7)
    bad = dict(pb.arrays)
    bad["vmapP"] = np.ascontiguousarray(np.roll(pb.arrays["vmapP"], 1, axis=2))
    pb2 = type(pb)(**{**pb.__dict__, "arrays": bad})
    rhs, _, _ = O.Oracle(pb2, "euler").rhs(u)
    w = pb.op_ref.M.sum(axis=0)
    tot = (pb.mesh.J[:, None] * (rhs[0].reshape(pb.K, pb.Np) * w)).sum()
    scale = (pb.mesh.J[:, None] * np.abs(rhs[0].reshape(pb.K, pb.Np)) * np.abs(w)).sum()
    assert abs(tot) > 1e-6 * scale


def test_advection_period_return():
    """v=(1,1,1) on the unit cube returns the field to itself at t=1; the DG
    error converges at high order between two resolutions."""
    errs = []
    for nhex in (3, 6):
        pb = cube_problem(3, nhex)
        u0 = ic_advection_sine(pb.arrays["px"], pb.arrays["py"], pb.arrays["pz"])
        u, _, t = O.Oracle(pb, "advection", cfl=0.5, tmax=1.0).run(u0, 10000)
        assert t == 1.0
        errs.append(l2_error(pb, _advection_exact(pb, 1.0), u[0]))
    assert errs[1] < errs[0] / 8  # at least O(h^3)


@pytest.mark.parametrize("dirichlet", [False, True])
@pytest.mark.parametrize("physics", PHYS)
def test_oracle_tangent_linear(physics, dirichlet):
    """The oracle's tangent-linear RHS (the reference for the GPU adjoint's
    dot-product test) agrees with central differences of the oracle RHS at
    frozen tau; for advection (linear in u) and periodic faces it equals the
    RHS of the direction itself.  With Dirichlet faces (x cut open, a fixed
    prescribed ub there) the face state does not move with u."""
    bc = {"x": "bc1", "y": "pbc", "z": "pbc"} if dirichlet else None
    pb = cube_problem(3, 2, bc=bc)
    u = smooth_state(pb, physics, seed=41)
    rng = np.random.default_rng(3)
    du = rng.standard_normal(u.shape)
    o = O.Oracle(pb, physics)
    if dirichlet:
        o.set_dirichlet(u[:, pb.mesh.vmapM.ravel()] * (1 + 0.1 * rng.standard_normal((u.shape[0], 1))))
    tl = O.rhs_tl(o, u, du)
    _, _, tau = o.rhs(u)
    if physics == "advection" and not dirichlet:
        assert rel_err(tl, O.rhs_tau(o, du, tau)) < 1e-14
    h = 1e-5 * np.abs(u).max() / np.abs(du).max()
    fd = (O.rhs_tau(o, u + h * du, tau) - O.rhs_tau(o, u - h * du, tau)) / (2 * h)
    assert rel_err(tl, fd) < 1e-8


def test_oracle_tangent_linear_catches_planted_error():
    """A dropped pressure term in a restated flux Jacobian is visible to the
    FD check above: perturbing gamma in the TL alone moves it by >> 1e-8."""
    pb = cube_problem(3, 2)
    u = smooth_state(pb, "euler", seed=41)
    du = np.random.default_rng(3).standard_normal(u.shape)
    tl = O.rhs_tl(O.Oracle(pb, "euler"), u, du)
    tl_bad = O.rhs_tl(O.Oracle(pb, "euler", gamma=1.4 * (1 + 1e-6)), u, du)
    assert rel_err(tl_bad, tl) > 1e-8


def test_trajectory_tangent_linear_vs_fd():
    """The CPU tangent-linear of a whole Euler LSRK trajectory (the tau(u) and
    dt(u) paths included) against central differences of the oracle
    forward, on a state without wave-speed ties; with tau and dt frozen it
    matches differences of the fixed-coefficient forward instead, and then
    misses the full derivative by far more than the tolerance (so dropping
    the tau/dt paths would fail the first check)."""
    from traj_tl import Trajectory

    from helpers import noisy_state

    pb = cube_problem(3, 2)
    u0 = noisy_state(pb, "euler", seed=3)
    rng = np.random.default_rng(5)
    wts = rng.uniform(0.5, 1.5, u0.shape)
    T = Trajectory(pb, "euler", cfl=0.4)
    n = 8
    uT, _ = T.run(u0, n)
    assert np.array_equal(uT, O.Oracle(pb, "euler", cfl=0.4).run(u0, n)[0])
    coeffs = list(T.coeffs)

    def J(x, fixed=None):
        return 0.5 * float(np.sum(wts * T.run(x, n, coeffs=fixed)[0] ** 2))

    h = 1e-6
    for q in range(2):
        d = rng.standard_normal(u0.shape) * np.abs(u0)
        uT, duT = T.run(u0, n, d)
        tl = float(np.sum(wts * uT * duT))
        fd = (J(u0 + h * d) - J(u0 - h * d)) / (2 * h)
        assert abs(tl - fd) <= 1e-6 * abs(fd), (tl, fd)
        uT, duT = T.run(u0, n, d, frozen=True)
        tlf = float(np.sum(wts * uT * duT))
        fdf = (J(u0 + h * d, coeffs) - J(u0 - h * d, coeffs)) / (2 * h)
        assert abs(tlf - fdf) <= 1e-6 * abs(fdf), (tlf, fdf)
        assert abs(tlf - fd) > 1e-4 * abs(fd)


@pytest.mark.parametrize("physics", PHYS)
def test_dirichlet_faces_reproduce_periodic_rhs(physics):
    """Dirichlet faces (EtoB == 1; the reference leaves the branch as
    "ub = function of x,y,z,t") take the prescribed face state ub.  Cut the
    periodic box open in x (tag bc1) and prescribe the periodic face average
    (u- + u+)/2 on the cut faces: with the periodic tau, the RHS is bitwise the
    periodic one; without the prescription it is not."""
    per = cube_problem(3, 3)
    cut = cube_problem(3, 3, bc={"x": "bc1", "y": "pbc", "z": "pbc"})
    assert (cut.arrays["EtoB"] == 1).sum() == 2 * 3 * 3 * 2
    u = noisy_state(per, physics, seed=7, eps=0.05)   # traces jump at every face
    nv = u.shape[0]
    o_per = O.Oracle(per, physics)
    r_per, _, tau_per = o_per.rhs(u)
    vmM = per.mesh.vmapM.ravel()
    g = 0.5 * (u[:, vmM] + u[:, per.arrays["vmapP"].ravel()])
    o_cut = O.Oracle(cut, physics)
    r_open = O.rhs_tau(o_cut, u, tau_per)
    assert np.abs(r_open - r_per).max() > 1e-6
    o_cut.set_dirichlet(g)
    np.testing.assert_array_equal(O.rhs_tau(o_cut, u, tau_per), r_per)
    assert nv == (5 if physics == "euler" else 1)


@pytest.mark.parametrize("physics", PHYS)
def test_dirichlet_free_stream(physics):
    """A free stream prescribed on Dirichlet faces (x and z cut open) stays a
    free stream: zero RHS."""
    pb = cube_problem(3, 3, bc={"x": "bc1", "y": "pbc", "z": "bc1"})
    KNp, KNF = pb.K * pb.Np, pb.K * 4 * pb.Nfp
    if physics == "euler":
        v = (0.3, -0.2, 0.1)
        s = np.array([1.0, *v, 1.0 / 0.4 + 0.5 * sum(x * x for x in v)])
    else:
        s = np.array([0.7])
    u = np.repeat(s[:, None], KNp, axis=1)
    o = O.Oracle(pb, physics)
    o.set_dirichlet(np.repeat(s[:, None], KNF, axis=1))
    r, _, _ = o.rhs(u)
    assert np.abs(r).max() < 1e-12


def test_dirichlet_exact_data_keeps_periodic_accuracy():
    """Euler density wave (the reference initial condition) on a box cut open
    in x, with the exact advected solution prescribed on the cut faces at the
    start of every step: after t = 0.05 the nodal error is no larger than
    the periodic box's (it is 5% smaller: the cut faces see exact data), and
    prescribing stale data (the state at t = 0) makes it four times larger."""
    from exadglm_amd.problem import uniform_problem

    def exact(x, y, z, t):
        return ic_euler_density_wave(x - t, y + 0.5 * t, z - t)

    tmax = 0.05
    errs = {}
    for name, bc, lag in (("per", None, None), ("cut", {"x": "bc1", "y": "pbc", "z": "pbc"}, False),
                          ("stale", {"x": "bc1", "y": "pbc", "z": "pbc"}, True)):
        pb = uniform_problem(3, 7, 7, 7, box=(0, 1, 0, 1, 0, 1), bc=bc)
        a = pb.arrays
        o = O.Oracle(pb, "euler", cfl=0.8, tmax=tmax)
        vm = pb.mesh.vmapM.ravel()
        fx, fy, fz = a["px"].ravel()[vm], a["py"].ravel()[vm], a["pz"].ravel()[vm]
        u, t = ic_euler_density_wave(a["px"], a["py"], a["pz"]), 0.0
        while t < tmax:
            if bc:
                o.set_dirichlet(exact(fx, fy, fz, 0.0 if lag else t))
            u, _, t = o.run(u, 1, t=t)
        ex = exact(a["px"].ravel(), a["py"].ravel(), a["pz"].ravel(), tmax)
        errs[name] = np.linalg.norm(u - ex)
    assert errs["cut"] <= errs["per"]
    assert errs["stale"] > 3 * errs["per"]


@pytest.mark.parametrize("N", [3, 4])
@pytest.mark.parametrize("physics", PHYS)
def test_oracle_rusanov_matches_textbook(physics, N):
    """The oracle's flux option (Rusanov for Euler, exact upwind for
    advection; the configurations' fluxes, which the reference lacks) on a
    state whose traces jump at every face: the RHS difference to the oracle's
    LM flux equals -LIFT (F*_rus - F*_lm) Fscale with both numerical fluxes
    restated in numpy from the textbook formulas.  The LM side is the
    reference's own flux (pinned against the compiled reference above), so
    this pins the Rusanov / upwind leg."""
    from helpers import lift_surface, textbook_surface_flux

    pb = cube_problem(N, 3)
    vel = (1.0, 0.5, -0.25)
    if physics == "euler":
        u = noisy_state(pb, "euler", seed=12 + N, eps=0.05)
    else:
        u = np.random.default_rng(8 + N).standard_normal((1, pb.K * pb.Np))
    r_lm, _, tau = O.Oracle(pb, physics, velocity=vel).rhs(u)
    r_rus, _, tau2 = O.Oracle(pb, physics, velocity=vel, flux="rusanov").rhs(u)
    assert np.array_equal(tau, tau2)
    f_rus = textbook_surface_flux(pb, u, physics, "rusanov", vel=vel)
    f_lm = textbook_surface_flux(pb, u, physics, "lm", tau=tau, vel=vel)
    expect = lift_surface(pb, f_rus - f_lm)
    diff = r_rus - r_lm
    assert np.abs(diff - expect).max() < 1e-12 * np.abs(expect).max()
    # the check sees a plausible mistake: the dissipation term with the wrong
    # sign (Euler) / the downwind trace (advection) misses by far more
    wrong = textbook_surface_flux(pb, u, physics, "rusanov", vel=vel, sign=-1.0)
    bad = lift_surface(pb, wrong - f_lm)
    assert np.abs(diff - bad).max() > 1e-3 * np.abs(expect).max()


@pytest.mark.parametrize("physics", PHYS)
def test_oracle_rusanov_equals_lm_on_continuous_traces(physics):
    """Where u- = u+ (a continuous nodal field: shared face nodes coincide)
    both fluxes reduce to n.F(u): the oracle RHS are equal to rounding."""
    pb = cube_problem(3, 3)
    u = smooth_state(pb, physics, seed=5)
    r_lm = O.Oracle(pb, physics).rhs(u)[0]
    r_rus = O.Oracle(pb, physics, flux="rusanov").rhs(u)[0]
    assert rel_err(r_rus, r_lm) < 1e-13


@pytest.mark.parametrize("physics", PHYS)
def test_oracle_rusanov_conserves(physics):
    """Flux telescoping with the Rusanov / upwind option: sum_e J_e 1^T M rhs_e
    vanishes on a periodic mesh for a discontinuous state too."""
    pb = cube_problem(3, 3)
    u = noisy_state(pb, physics, seed=9, eps=0.05)
    rhs, _, _ = O.Oracle(pb, physics, flux="rusanov").rhs(u)
    w = pb.op_ref.M.sum(axis=0)
    for v in range(rhs.shape[0]):
        tot = (pb.mesh.J[:, None] * (rhs[v].reshape(pb.K, pb.Np) * w)).sum()
        scale = (pb.mesh.J[:, None] * np.abs(rhs[v].reshape(pb.K, pb.Np)) * np.abs(w)).sum()
        assert abs(tot) < 1e-13 * scale


def test_oracle_tangent_linear_refuses_rusanov():
    pb = cube_problem(2, 2)
    u = smooth_state(pb, "euler", seed=1)
    with pytest.raises(ValueError):
        O.rhs_tl(O.Oracle(pb, "euler", flux="rusanov"), u, u)
