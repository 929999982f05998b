"""Problem setup: uniform periodic meshes with operators, and initial conditions.

`uniform_problem` assembles exactly the array set that the reference's
save_to_binaries writes to datain/ (dglm_3d_launcher.py:74-149), built from
TetraRef/Operator3DRef/Operator3DFace and Uniform3DSingle/Mesh3D.  The
initial conditions are the reference's two cases (advec_init.cpp:5-23,
euler_init.cpp:5-35) plus the isentropic vortex, random-Fourier and
Gaussian-bump fields used by the synthetic workloads.
"""
import math
from dataclasses import dataclass, field

import numpy as np

from .mesh import Mesh3D, Uniform3DSingle
from .reference import Operator3DFace, Operator3DRef, TetraRef

DATAIN_NAMES = ["EtoB", "vmapP", "vmapF", "Fmask", "WDr", "WDs", "WDt", "LIFT", "Fscale", "rx", "ry", "rz", "sx",
                "sy", "sz", "tx", "ty", "tz", "nx", "ny", "nz", "px", "py", "pz"]


@dataclass
class Problem:
    N: int
    K: int
    Np: int
    Nfp: int
    tet: TetraRef
    op_ref: Operator3DRef
    op_face: Operator3DFace
    mesh: Mesh3D
    arrays: dict = field(default_factory=dict)  # datain name -> contiguous array


def uniform_problem(N, nx, ny, nz, box=(0, 1, 0, 1, 0, 1), bc=None, interpN=2):
    """Mesh of (nx-1)(ny-1)(nz-1) hexes x 6 tets on box = (x1,x2,y1,y2,z1,z2)."""
    tet = TetraRef(N, interpN)
    op_ref = Operator3DRef(tet)
    op_face = Operator3DFace(op_ref)
    elem = Uniform3DSingle(nx, ny, nz, *box, bc=bc)
    mesh = Mesh3D(op_face, elem)
    c = np.ascontiguousarray
    arr = {
        "EtoB": c(mesh.EtoB, np.int32), "vmapP": c(mesh.vmapP, np.int32), "vmapF": c(mesh.vmapF, np.int32),
        "Fmask": c(op_face.Fmask, np.int32),
        "WDr": c(op_ref.WDr), "WDs": c(op_ref.WDs), "WDt": c(op_ref.WDt), "LIFT": c(op_face.LIFT),
        "Fscale": c(mesh.Fscale), "rx": c(mesh.rx), "ry": c(mesh.ry), "rz": c(mesh.rz), "sx": c(mesh.sx),
        "sy": c(mesh.sy), "sz": c(mesh.sz), "tx": c(mesh.tx), "ty": c(mesh.ty), "tz": c(mesh.tz),
        "nx": c(mesh.NX), "ny": c(mesh.NY), "nz": c(mesh.NZ), "px": c(mesh.PX), "py": c(mesh.PY), "pz": c(mesh.PZ),
    }
    return Problem(N=N, K=mesh.Nelem, Np=tet.Np, Nfp=tet.Nfp, tet=tet, op_ref=op_ref, op_face=op_face, mesh=mesh,
                   arrays=arr)


def cube_problem(N, nhex, box=(0, 1, 0, 1, 0, 1), bc=None):
    """nhex^3 hexes (the reference's nx = nhex + 1) x 6 tets."""
    return uniform_problem(N, nhex + 1, nhex + 1, nhex + 1, box, bc)


# ---------------------------------------------------------------------------
# Initial conditions.  All return float64 arrays shaped [nvar, K*Np].
# ---------------------------------------------------------------------------
def ic_advection_sine(px, py, pz):
    """u0 = sin(2 pi x) sin(2 pi y) sin(2 pi z) (advec_init.cpp:18-21)."""
    x, y, z = px.ravel(), py.ravel(), pz.ravel()
    return (np.sin(2 * np.pi * x) * np.sin(2 * np.pi * y) * np.sin(2 * np.pi * z))[None, :].copy()


def ic_euler_density_wave(px, py, pz, gamma=1.4):
    """rho = 1 + 0.2 sin(2 pi (x+y+z)), v = (1, -0.5, 1), p = 1 (euler_init.cpp:11-33)."""
    x, y, z = px.ravel(), py.ravel(), pz.ravel()
    vx, vy, vz, pres = 1.0, -0.5, 1.0, 1.0
    u1 = 1 + 0.2 * np.sin(2 * np.pi * (x + y + z))
    return np.stack([u1, u1 * vx, u1 * vy, u1 * vz, pres / (gamma - 1) + 0.5 * u1 * (vx * vx + vy * vy + vz * vz)])


def isentropic_vortex(x, y, z, t=0.0, gamma=1.4, beta=5.0, x0=5.0, y0=5.0, u_inf=(1.0, 1.0, 0.0),
                      period=(10.0, 10.0), images=1):
    """Conserved variables of the isentropic vortex (Shu 1998) advected by u_inf.

    rho = T^(1/(g-1)), T = 1 - (g-1) beta^2 / (8 g pi^2) exp(1-r^2),
    (u,v) = u_inf + beta/(2 pi) exp((1-r^2)/2) (-(y-yc), x-xc), p = rho^g.
    On a periodic box the perturbations of the (2 images+1)^2 nearest copies
    are superposed (the tail at r = 5 is ~6e-6 of the peak).
    """
    x, y, z = np.ravel(x), np.ravel(y), np.ravel(z)
    xc = x0 + u_inf[0] * t
    yc = y0 + u_inf[1] * t
    du = np.zeros_like(x)
    dv = np.zeros_like(x)
    dT = np.zeros_like(x)
    for ix in range(-images, images + 1):
        for iy in range(-images, images + 1):
            # centre of this periodic copy, wrapped into the box first
            cx = (xc % period[0]) + ix * period[0]
            cy = (yc % period[1]) + iy * period[1]
            dx, dy = x - cx, y - cy
            r2 = dx * dx + dy * dy
            ex = np.exp(0.5 * (1 - r2))
            du += -beta / (2 * np.pi) * dy * ex
            dv += beta / (2 * np.pi) * dx * ex
            dT += -(gamma - 1) * beta * beta / (8 * gamma * np.pi * np.pi) * ex * ex
    T = 1 + dT
    rho = T ** (1 / (gamma - 1))
    u = u_inf[0] + du
    v = u_inf[1] + dv
    w = np.full_like(x, u_inf[2])
    p = rho ** gamma
    E = p / (gamma - 1) + 0.5 * rho * (u * u + v * v + w * w)
    return np.stack([rho, rho * u, rho * v, rho * w, E])


def fourier_modes(seed, nmodes=8, kmax=4):
    """Random Fourier field spec: k uniform over {k in Z^3: 0<|k|<=kmax} (listed
    in lexicographic order of (kx,ky,kz)), a ~ N(0,1)/|k|, phase ~ U[0,2pi),
    drawn in that order from numpy.random.default_rng(seed)."""
    r = range(-kmax, kmax + 1)
    ks = np.array([(a, b, c) for a in r for b in r for c in r if 0 < a * a + b * b + c * c <= kmax * kmax])
    rng = np.random.default_rng(seed)
    idx = rng.integers(0, len(ks), nmodes)
    k = ks[idx].astype(np.float64)
    kn = np.sqrt((k * k).sum(axis=1))
    amp = rng.standard_normal(nmodes) / kn
    phase = rng.uniform(0.0, 2 * np.pi, nmodes)
    return k, amp, phase


def ic_fourier(px, py, pz, seed, nmodes=8, kmax=4, shift=(0.0, 0.0, 0.0)):
    """u = sum_m a_m cos(2 pi k_m.(x - shift) + phi_m) on the unit-periodic cube."""
    k, amp, phase = fourier_modes(seed, nmodes, kmax)
    x, y, z = px.ravel() - shift[0], py.ravel() - shift[1], pz.ravel() - shift[2]
    u = np.zeros_like(x)
    for m in range(nmodes):
        u += amp[m] * np.cos(2 * np.pi * (k[m, 0] * x + k[m, 1] * y + k[m, 2] * z) + phase[m])
    return u[None, :]


def gaussian_bumps(px, py, pz, seed, box, nbumps=4, width=0.5, amplitude=1e-3):
    """amplitude * sum_g N(0,1) exp(-|x-c_g|^2 / (2 width^2)), centres uniform in
    the box, minimum-image distance on the periodic box; draws from
    default_rng(seed): centres (nbumps x 3 uniforms) then weights."""
    rng = np.random.default_rng(seed)
    lo = np.array([box[0], box[2], box[4]])
    hi = np.array([box[1], box[3], box[5]])
    L = hi - lo
    cen = lo + rng.uniform(size=(nbumps, 3)) * L
    w = rng.standard_normal(nbumps)
    x = np.stack([px.ravel(), py.ravel(), pz.ravel()], axis=1)
    out = np.zeros(x.shape[0])
    for g in range(nbumps):
        d = x - cen[g]
        d -= L * np.round(d / L)
        out += w[g] * np.exp(-(d * d).sum(axis=1) / (2 * width * width))
    return amplitude * out


def perturb_rho_p(u, bump, gamma=1.4):
    """Add `bump` to density and pressure of conserved state u, keeping velocity."""
    rho, mx, my, mz, E = u
    vx, vy, vz = mx / rho, my / rho, mz / rho
    p = (gamma - 1) * (E - 0.5 * rho * (vx * vx + vy * vy + vz * vz))
    rho2 = rho + bump
    p2 = p + bump
    return np.stack([rho2, rho2 * vx, rho2 * vy, rho2 * vz,
                     p2 / (gamma - 1) + 0.5 * rho2 * (vx * vx + vy * vy + vz * vz)])


def vortex_box(nhex):
    """Domain of the vortex workloads: [0,10]^2 x [0,10/3] with nhex^3 hexes."""
    return (0.0, 10.0, 0.0, 10.0, 0.0, 10.0 / 3.0)


def l2_error(problem, ref_sol, numeric_sol, apply_sqrt=True):
    """sqrt(sum_e J_e E_e^T M_high E_e) with E = u_ref(high nodes) - interpM u_e
    (dglm_3d_launcher.py:328-355)."""
    K, Np, hNp = problem.K, problem.Np, problem.tet.high_Np
    u = np.asarray(numeric_sol, dtype=np.float64).reshape(K, Np)
    r = np.asarray(ref_sol, dtype=np.float64).reshape(K, hNp)
    E = r - u @ problem.op_ref.interpM.T
    Eint = problem.mesh.J * np.einsum("ei,ij,ej->e", E, problem.op_ref.high_M, E)
    s = np.sum(Eint)
    return np.sqrt(s) if apply_sqrt else s
