"""Shared builders for the tests: small problems and smooth random states."""
import numpy as np

from exadglm_amd.problem import cube_problem, ic_fourier


def smooth_state(pb, physics, seed=0, gamma=1.4):
    """A smooth, physically valid state on the unit-periodic box."""
    a = pb.arrays
    px, py, pz = a["px"], a["py"], a["pz"]
    if physics == "advection":
        return ic_fourier(px, py, pz, seed, nmodes=6, kmax=2)
    f = [ic_fourier(px, py, pz, seed + i, nmodes=4, kmax=2)[0] for i in range(5)]
    rho = 1.0 + 0.1 * f[0]
    vx, vy, vz = 0.5 + 0.2 * f[1], -0.3 + 0.2 * f[2], 0.2 * f[3]
    p = 1.0 + 0.1 * f[4]
    E = p / (gamma - 1) + 0.5 * rho * (vx * vx + vy * vy + vz * vz)
    return np.stack([rho, rho * vx, rho * vy, rho * vz, E])


def noisy_state(pb, physics, seed=0, eps=1e-2):
    """smooth_state with independent relative noise eps per node: the traces
    on the two sides of a face differ, so the wave-speed maxima (per face,
    and the global one that sets dt) are attained at single nodes and faces.
    On a continuous state the two sides of every face carry the same nodes,
    the global maximum of lambda*Fscale is a tie between a face and its
    neighbour, and dt(u) is not differentiable there."""
    u = smooth_state(pb, physics, seed)
    return u * (1 + eps * np.random.default_rng(seed + 1000).standard_normal(u.shape))


def rel_err(a, b):
    return float(np.linalg.norm(np.ravel(a - b)) / max(np.linalg.norm(np.ravel(b)), 1e-300))


def rel_err_per_var(a, b):
    a = np.asarray(a).reshape(np.shape(b)[0], -1)
    b = np.asarray(b).reshape(a.shape)
    return [rel_err(a[v], b[v]) for v in range(a.shape[0])]


def velocity_for(pb, v):
    n = pb.K * pb.Np
    return [np.full(n, float(c)) for c in v]
