"""CPU tangent-linear of the whole LSRK trajectory, built from oracle pieces
(test infrastructure).  Forward mode, including the derivatives of the wave
speed (argmax nodes as calc_maxv picks them), of tau and of dt; with
`frozen=True` tau and dt are held at their forward values.

It is pinned by central differences of the oracle forward
(tests/test_oracle.py::test_trajectory_tangent_linear_vs_fd) and is the
reference for exact dot-product tests of the GPU adjoint."""
import numpy as np

import pyoracle as O

RKA = [0.0, -567301805773.0 / 1357537059087.0, -2404267990393.0 / 2016746695238.0,
       -3550918686646.0 / 2091501179385.0, -1275806237668.0 / 842570457699.0]
RKB = [1432997174477.0 / 9575080441755.0, 5161836677717.0 / 13612068292357.0,
       1720146321549.0 / 2090206949498.0, 3134564353537.0 / 4481467310338.0,
       2277821191437.0 / 14882151754819.0]


class Trajectory:
    def __init__(self, pb, physics, cfl=0.4, gamma=1.4, tau_scale=1.2, velocity=None):
        self.pb, self.physics = pb, physics
        self.o = O.Oracle(pb, physics, cfl=cfl, gamma=gamma, tau_scale=tau_scale, velocity=velocity)
        self.cfl, self.g, self.ts = cfl, gamma, tau_scale
        self.nv = 5 if physics == "euler" else 1
        a = pb.arrays
        self.K, self.Np, self.Nfp, self.N = pb.K, pb.Np, pb.Nfp, pb.N
        self.Fm = a["Fmask"].reshape(4, self.Nfp)
        self.fs = a["Fscale"].reshape(self.K, 4)
        self.vF = a["vmapF"].reshape(self.K, 4)
        if velocity is None:
            velocity = (1.0, 1.0, 1.0)
        n = self.K * self.Np
        self.vel = np.stack([np.broadcast_to(np.asarray(c, dtype=np.float64), (n,)) for c in velocity])

    # wave speed per face and d(lambda) per face for a state direction
    def _face_vals(self, u):
        K, Np = self.K, self.Np
        idx = (np.arange(K)[:, None, None] * Np + self.Fm[None, :, :])  # K,4,Nfp
        if self.physics == "euler":
            U = u[:, idx]  # 5,K,4,Nfp
            v = U[1:4] / U[0]
            v2 = (v * v).sum(0)
            pres = (self.g - 1) * (U[4] - 0.5 * U[0] * v2)
            c2 = np.abs(self.g * pres / U[0])
            return idx, U, v2, c2
        V = self.vel[:, idx]
        return idx, None, (V * V).sum(0), None

    def lam(self, u):
        idx, U, v2, c2 = self._face_vals(u)
        lam = np.sqrt(v2.max(-1))
        if c2 is not None:
            lam = lam + np.sqrt(c2.max(-1))
        return lam

    def dlam(self, u, du):
        if self.physics != "euler":
            return np.zeros((self.K, 4))
        idx, U, v2, c2 = self._face_vals(u)
        dU = du[:, idx]
        g = self.g
        out = np.zeros((self.K, 4))
        kk, ff = np.meshgrid(np.arange(self.K), np.arange(4), indexing="ij")
        for vals, kind in ((v2, "v"), (c2, "c")):
            j = vals.argmax(-1)
            mx = vals.max(-1)
            Uj = U[:, kk, ff, j]
            dj = dU[:, kk, ff, j]
            r, m, E = Uj[0], Uj[1:4], Uj[4]
            vv = m / r
            q2 = (vv * vv).sum(0)
            if kind == "v":
                d = (-2 * q2 / r) * dj[0] + (2 * vv / r * dj[1:4]).sum(0)
            else:
                p = (g - 1) * (E - 0.5 * r * q2)
                sg = np.where(g * p / r >= 0, 1.0, -1.0)
                d = sg * g * ((g - 1) * 0.5 * q2 / r - p / r ** 2) * dj[0] + \
                    (sg * g * (-(g - 1) * vv / r) * dj[1:4]).sum(0) + sg * g * (g - 1) / r * dj[4]
            out += np.where(mx > 0, d / (2 * np.sqrt(np.where(mx > 0, mx, 1.0))), 0.0)
        return out

    def tau(self, lam):
        return 0.25 * (lam + lam.ravel()[self.vF.ravel()].reshape(self.K, 4)) * self.ts

    def run(self, u0, nsteps, du0=None, frozen=False, coeffs=None):
        """Forward trajectory; with du0, also the tangent du(T).  The (tau, dt)
        of every step are kept in self.coeffs; passing such a list as
        `coeffs` runs the forward with those coefficients held fixed (the
        function whose derivative the frozen mode is)."""
        u = np.array(u0, dtype=np.float64).reshape(self.nv, -1)
        self.coeffs = []
        du = None if du0 is None else np.array(du0, dtype=np.float64).reshape(self.nv, -1)
        k = np.zeros_like(u)
        dk = None if du is None else np.zeros_like(u)
        cf = self.cfl / (2 * self.N + 1)
        for st in range(nsteps):
            lam = self.lam(u)
            lfs = lam * self.fs
            fstar = int(np.argmax(lfs.ravel()))
            dt = cf * (1 / lfs.ravel()[fstar])
            tau = self.tau(lam)
            if coeffs is not None:
                tau, dt = coeffs[st]
            self.coeffs.append((tau, dt))
            if du is not None:
                if frozen:
                    ddt, dtau = 0.0, np.zeros_like(tau)
                else:
                    dl = self.dlam(u, du)
                    ddt = -dt / lam.ravel()[fstar] * dl.ravel()[fstar]
                    dtau = self.tau(dl)
            for s in range(5):
                r = O.rhs_tau(self.o, u, tau)
                if du is not None:
                    dr = O.rhs_tl(self.o, u, du, tau)
                    if np.any(dtau != 0):
                        c = np.abs(tau).max() / np.abs(dtau).max()
                        dr = dr + (O.rhs_tau(self.o, u, tau + c * dtau) - r) / c
                    dk = RKA[s] * dk + ddt * r + dt * dr
                    du = du + RKB[s] * dk
                k = RKA[s] * k + dt * r
                u = u + RKB[s] * k
        return u, du
