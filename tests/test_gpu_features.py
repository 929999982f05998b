"""hipGraph replay of the step loop, device conservation sums, and the
Rusanov / upwind flux option."""
import math

import numpy as np
import pytest

from helpers import noisy_state, rel_err, rel_err_per_var, smooth_state
from exadglm_amd.problem import cube_problem, ic_advection_sine, isentropic_vortex, uniform_problem, vortex_box
from exadglm_amd.solver import Solver
import pyoracle as O

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


@pytest.mark.parametrize("physics,precision", [("advection", "fp64"), ("euler", "fp64"), ("euler", "fp32")])
def test_graph_replay_matches_eager(gpu, physics, precision):
    """Replayed two-step graphs give bitwise the eager result, over an odd
    number of steps, across repeated calls and with tmax clipping."""
    _say(f"graph replay {physics} {precision}")
    pb = cube_problem(3, 4)
    u0 = smooth_state(pb, physics, seed=71)
    out = []
    for graphs in (False, True):
        with Solver(pb, physics, cfl=0.5, tmax=0.05, precision=precision) as s:
            s.set_graphs(graphs)
            s.set_state(u0)
            s.step(7)
            a = s.get_state()
            s.step(9)
            b = s.get_state()
            s.step(400)  # runs into tmax: the remaining steps are no-ops
            t, _, ts, done = s.time()
            c = s.get_state()
            out.append((a, b, c, t, ts, done))
    for x, y in zip(out[0][:3], out[1][:3]):
        assert np.array_equal(x, y)
    assert out[0][3:] == out[1][3:] and out[1][5] and out[1][3] == 0.05


def test_device_integrals_match_host(gpu):
    _say("device integrals")
    pb = cube_problem(4, 3)
    u = smooth_state(pb, "euler", seed=3)
    with Solver(pb, "euler") as s:
        s.set_state(u)
        I = s.integrals(pb.mesh.J, pb.op_ref.M)
    w = pb.op_ref.M.sum(axis=0)
    ref = [float((pb.mesh.J[:, None] * (u[v].reshape(pb.K, pb.Np) * w)).sum()) for v in range(5)]
    assert rel_err(np.array(I), np.array(ref)) < 1e-14


@pytest.mark.parametrize("physics", ["advection", "euler"])
@pytest.mark.parametrize("variant", ["simple", "mfma", "mfma_ws"])
def test_rusanov_equals_lm_on_continuous_traces(gpu, physics, variant):
    """Where u- = u+ (nodal interpolant of a continuous field: shared face
    nodes coincide) both fluxes reduce to n.F(u): the RHS are equal."""
    if variant == "mfma_ws" and physics != "euler":
        pytest.skip("the warp-specialised kernel is Euler only")
    _say(f"rusanov continuous {physics} {variant}")
    pb = cube_problem(3, 3)
    u = smooth_state(pb, physics, seed=5)
    out = []
    for flux in ("lm", "rusanov"):
        with Solver(pb, physics, flux=flux) as s:
            s.set_stage_kernel(variant)
            out.append(s.rhs(u)[0])
    assert rel_err(out[1], out[0]) < 1e-13


@pytest.mark.parametrize("variant", ["simple", "mfma"])
def test_upwind_advection_surface_term(gpu, variant):
    """Advection with the Rusanov option is exact upwinding: on a
    discontinuous field the RHS difference to the LM flux equals
    -LIFT (F*_upwind - F*_LM) Fscale computed from the traces."""
    _say(f"upwind advection {variant}")
    pb = cube_problem(3, 3)
    rng = np.random.default_rng(8)
    u = rng.standard_normal((1, pb.K * pb.Np))  # discontinuous across faces
    vel = (1.0, 0.5, -0.25)
    res = {}
    for flux in ("lm", "rusanov"):
        with Solver(pb, "advection", flux=flux) as s:
            s.set_stage_kernel(variant)
            s.set_velocity(*vel)
            res[flux] = s.rhs(u)
    _, _, tau = res["lm"]
    a = pb.arrays
    K, Np, Nfp = pb.K, pb.Np, pb.Nfp
    uu = u[0]
    Fm = a["Fmask"].reshape(4, Nfp)
    uM = uu.reshape(K, Np)[:, Fm.ravel()].reshape(K, 4, Nfp)
    uP = uu[a["vmapP"].reshape(K, 4, Nfp)]
    nvec = np.stack([a["nx"], a["ny"], a["nz"]], -1)
    vn = (nvec @ np.array(vel))[:, :, None]
    ub = 0.5 * (uM + uP)
    f_up = np.where(vn > 0, vn * uM, vn * uP)
    f_lm = vn * ub - tau[:, :, None] * (ub - uM)
    dfl = ((f_up - f_lm) * a["Fscale"][:, :, None]).reshape(K, 4 * Nfp)
    expect = -(dfl @ a["LIFT"].T).reshape(1, -1)
    diff = res["rusanov"][0] - res["lm"][0]
    assert np.abs(diff - expect).max() < 1e-12 * np.abs(expect).max()


def test_rusanov_euler_vortex_runs_and_conserves(gpu):
    _say("rusanov euler conservation")
    pb = uniform_problem(4, 7, 7, 7, box=vortex_box(6))
    a = pb.arrays
    u0 = isentropic_vortex(a["px"], a["py"], a["pz"])
    with Solver(pb, "euler", cfl=0.4, flux="rusanov") as s:
        s.set_state(u0)
        I0 = s.integrals(pb.mesh.J, pb.op_ref.M)
        s.step(30)
        I1 = s.integrals(pb.mesh.J, pb.op_ref.M)
        u = s.get_state()
    assert np.isfinite(u).all()
    for v in (0, 1, 2, 4):
        assert abs(I1[v] - I0[v]) <= 1e-12 * abs(I0[v])


@pytest.mark.parametrize("N", [3, 4])
@pytest.mark.parametrize("variant", ["simple", "mfma", "mfma_ws"])
def test_rusanov_euler_surface_term(gpu, variant, N):
    """Euler with the Rusanov option on a discontinuous state: the RHS
    difference to the LM flux equals -LIFT (F*_Rusanov - F*_LM) Fscale with
    both fluxes computed from the traces in numpy (N = 4 is the order of the
    benchmark workload and its warp-specialised kernel)."""
    _say(f"rusanov euler surface term {variant} N={N}")
    pb = cube_problem(N, 3)
    from helpers import noisy_state

    u = noisy_state(pb, "euler", seed=12, eps=0.05)
    res = {}
    for flux in ("lm", "rusanov"):
        with Solver(pb, "euler", flux=flux) as s:
            s.set_stage_kernel(variant)
            res[flux] = s.rhs(u)
    _, _, tau = res["lm"]
    a = pb.arrays
    K, Np, Nfp = pb.K, pb.Np, pb.Nfp
    g = 1.4
    Fm = a["Fmask"].reshape(4, Nfp)
    uM = np.stack([u[v].reshape(K, Np)[:, Fm.ravel()].reshape(K, 4, Nfp) for v in range(5)])
    uP = np.stack([u[v][a["vmapP"].reshape(K, 4, Nfp)] for v in range(5)])
    n = np.stack([a["nx"], a["ny"], a["nz"]])[:, :, :, None]

    def nflux(U):
        inv = 1 / U[0]
        p = (g - 1) * (U[4] - 0.5 * (U[1] ** 2 + U[2] ** 2 + U[3] ** 2) * inv)
        vn = (n[0] * U[1] + n[1] * U[2] + n[2] * U[3]) * inv
        F = np.stack([vn * U[0], U[1] * vn + n[0] * p, U[2] * vn + n[1] * p, U[3] * vn + n[2] * p, (U[4] + p) * vn])
        return F, np.abs(vn) + np.sqrt(np.abs(g * p * inv))

    fm, lm = nflux(uM)
    fp, lp = nflux(uP)
    f_rus = 0.5 * (fm + fp) + 0.5 * np.maximum(lm, lp) * (uM - uP)
    ub = 0.5 * (uM + uP)
    fb, _ = nflux(ub)
    f_lm = fb - tau[None, :, :, None] * (ub - uM)
    dfl = ((f_rus - f_lm) * a["Fscale"][None, :, :, None]).reshape(5, K, 4 * Nfp)
    expect = -(dfl @ a["LIFT"].T).reshape(5, -1)
    diff = res["rusanov"][0] - res["lm"][0]
    assert np.abs(diff - expect).max() < 1e-12 * np.abs(expect).max()


_PER_ELEMENT = ("EtoB", "vmapP", "vmapF", "Fscale", "rx", "ry", "rz", "sx", "sy", "sz", "tx", "ty", "tz", "nx", "ny",
                "nz", "px", "py", "pz")


def _relabel(pb, seed):
    """The problem's datain arrays with its elements relabelled by a random
    permutation: new element e is old element old[e] (rows reordered, and the
    element parts of the vmapP node indices and vmapF face indices mapped)."""
    K, Np = pb.K, pb.Np
    old = np.random.default_rng(seed).permutation(K)
    new_of_old = np.empty(K, np.int64)
    new_of_old[old] = np.arange(K)
    arr = dict(pb.arrays)
    for k in _PER_ELEMENT:
        arr[k] = np.ascontiguousarray(pb.arrays[k][old])
    vp = arr["vmapP"].astype(np.int64)
    arr["vmapP"] = np.ascontiguousarray((new_of_old[vp // Np] * Np + vp % Np).astype(np.int32))
    vf = arr["vmapF"].astype(np.int64)
    arr["vmapF"] = np.ascontiguousarray((new_of_old[vf // 4] * 4 + vf % 4).astype(np.int32))
    return arr, old


@pytest.mark.parametrize("N,nhex", [(4, 6), (3, 7)])
def test_warp_specialised_kernel_is_independent_of_element_order(gpu, N, nhex):
    """The warp-specialised kernel visits batches in the Morton order of their
    node centroids; the order is a schedule only: a randomly relabelled mesh,
    and a mesh without node coordinates (mesh order), give bitwise the same
    states after 4 steps (K = 1296: whole batches; K = 2058: a short last
    batch)."""
    _say(f"ws element order N={N} nhex={nhex}")
    pb = cube_problem(N, nhex)
    K, Np = pb.K, pb.Np
    u0 = noisy_state(pb, "euler", seed=5)
    arr, old = _relabel(pb, 11)
    no_xyz = {k: v for k, v in pb.arrays.items() if k not in ("px", "py", "pz")}
    runs = []
    for arrays, u in ((pb.arrays, u0), (arr, np.ascontiguousarray(u0.reshape(5, K, Np)[:, old].reshape(5, -1))),
                      (no_xyz, u0)):
        with Solver(pb, "euler", cfl=0.4, tmax=math.inf, arrays=arrays, K=K, flux="rusanov") as s:
            s.set_stage_kernel("mfma_ws")
            assert s.stage_kernel() == "mfma_ws"
            s.set_state(u)
            s.step(4)
            runs.append(s.get_state())
    back = np.empty_like(runs[1]).reshape(5, K, Np)
    back[:, old] = runs[1].reshape(5, K, Np)
    assert np.isfinite(runs[0]).all()
    assert np.array_equal(runs[0], back.reshape(5, -1))
    assert np.array_equal(runs[0], runs[2])


@pytest.mark.parametrize("where", ["all", "one node"])
def test_non_finite_state_is_reported(gpu, where):
    """A state that turns non-finite is reported by dglm_get_time as
    DGLM_ERR_STATE: a NaN anywhere reaches the wave speed of some face within
    a step, the device-side dt becomes 0 and the step is flagged (a host-side
    read of the device flag; the kernels run normally on NaN data)."""
    from exadglm_amd import _capi as A

    _say(f"non-finite state ({where})")
    pb = cube_problem(3, 3)
    u = smooth_state(pb, "euler", seed=2)
    if where == "all":
        u[:] = np.nan
    else:
        u[0, 17] = np.nan  # an interior node of element 0
    with Solver(pb, "euler", cfl=0.4) as s:
        s.set_state(smooth_state(pb, "euler", seed=2))
        s.step(2)
        s.time()  # a finite state steps without complaint
        s.set_state(u)
        s.step(3)
        with pytest.raises(A.DGLMError) as ei:
            s.time()
        assert ei.value.code == A.DGLM_ERR_STATE
