"""The synthetic workloads of the benchmark suite (meshes, initial states,
objectives).  All meshes are uniform, periodic, 6 tetrahedra per hex.

  C1  advection N=3 on the unit cube, 10^3 hexes (K=6000), v=(1,1,1),
      u0 = sin(2 pi x) sin(2 pi y) sin(2 pi z), CFL 0.5, 100 steps
  C2  Euler N=4, isentropic vortex on [0,10]^2 x [0,10/3], 32^3 hexes
      (K=196,608), CFL 0.4, 200 steps, FP64
  C3  the same vortex, N=6 on 24^3 hexes (K=82,944), FP32-MFMA contraction,
      500 steps
  C4  advection adjoint, N=5 on 20^3 hexes (K=48,000), v=(1,0.5,0.25),
      u0 = random Fourier field (seed 0), target = the forward run of the
      seed-1 field, J = 1/2 ||u(T) - u_target||^2_L2, 400 steps, checkpoints
      every 20
  C5  Euler adjoint, vortex of C2 plus 1e-3 Gaussian bumps (seed 2) in rho
      and p, J = int 1/2 rho |v|^2 at T after 300 steps, gradient w.r.t. the
      bump field
"""
import math

import numpy as np

from .problem import (cube_problem, gaussian_bumps, ic_advection_sine, ic_fourier, isentropic_vortex,
                      uniform_problem, vortex_box)

GAMMA = 1.4


def c1():
    pb = cube_problem(3, 10)
    u0 = ic_advection_sine(pb.arrays["px"], pb.arrays["py"], pb.arrays["pz"])
    return dict(problem=pb, physics="advection", u0=u0, cfl=0.5, steps=100, velocity=(1.0, 1.0, 1.0))


def vortex(N, nhex):
    pb = uniform_problem(N, nhex + 1, nhex + 1, nhex + 1, box=vortex_box(nhex))
    a = pb.arrays
    return pb, isentropic_vortex(a["px"], a["py"], a["pz"], gamma=GAMMA)


def c2():
    pb, u0 = vortex(4, 32)
    return dict(problem=pb, physics="euler", u0=u0, cfl=0.4, steps=200)


def c3():
    pb, u0 = vortex(6, 24)
    return dict(problem=pb, physics="euler", u0=u0, cfl=0.4, steps=500, precision="fp32")


def c4(nhex=20, N=5):
    pb = cube_problem(N, nhex)
    a = pb.arrays
    return dict(problem=pb, physics="advection", u0=ic_fourier(a["px"], a["py"], a["pz"], 0),
                u_target0=ic_fourier(a["px"], a["py"], a["pz"], 1), cfl=0.5, steps=400, every=20,
                velocity=(1.0, 0.5, 0.25))


def c5(nhex=32, N=4):
    pb, base = vortex(N, nhex)
    a = pb.arrays
    bump = gaussian_bumps(a["px"], a["py"], a["pz"], 2, vortex_box(nhex))
    return dict(problem=pb, physics="euler", base=base, bump=bump, cfl=0.4, steps=300,
                every=int(round(math.sqrt(300))))


# ---------------------------------------------------------------------------
# Objectives and their gradients with respect to the final state
# ---------------------------------------------------------------------------
def mass_weights(pb):
    """Quadrature weights per node, J_e * (1^T M)_i, so that sum w*f = int f."""
    return (pb.mesh.J[:, None] * pb.op_ref.M.sum(axis=0)[None, :]).ravel()


def l2_misfit(pb, u, ut):
    """J = 1/2 int (u - ut)^2 (element mass matrices) and dJ/du."""
    K, Np = pb.K, pb.Np
    d = (u - ut).reshape(K, Np)
    Md = d @ pb.op_ref.M  # M symmetric
    J = 0.5 * float(np.sum(pb.mesh.J[:, None] * d * Md))
    return J, (pb.mesh.J[:, None] * Md).reshape(1, -1)


def kinetic_energy(pb, u):
    """J = int 1/2 |m|^2 / rho (nodal quadrature with the mass-matrix weights) and dJ/du."""
    w = mass_weights(pb)
    rho, m = u[0], u[1:4]
    ke = 0.5 * (m * m).sum(axis=0) / rho
    g = np.zeros_like(u)
    g[0] = -ke / rho * w
    g[1:4] = m / rho * w
    return float(np.sum(w * ke)), g


def perturbed_state(base, eps, gamma=GAMMA):
    """base vortex with eps added to rho and p (velocity kept), and d(u0)/d(eps)
    per node (for the chain rule to the perturbation field)."""
    rho, mx, my, mz, E = base
    vx, vy, vz = mx / rho, my / rho, mz / rho
    q2 = vx * vx + vy * vy + vz * vz
    p = (gamma - 1) * (E - 0.5 * rho * q2)
    r2, p2 = rho + eps, p + eps
    u0 = np.stack([r2, r2 * vx, r2 * vy, r2 * vz, p2 / (gamma - 1) + 0.5 * r2 * q2])
    du_deps = np.stack([np.ones_like(rho), vx, vy, vz, 1 / (gamma - 1) + 0.5 * q2])
    return u0, du_deps
