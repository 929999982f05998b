"""Reference-element objects with the reference's Python API.

TetraRef, Operator3DRef and Operator3DFace mirror dim3/src/python/tetra.py:305-331,
operator_3d_ref.py:302-349 and operator_3d_face.py:275-301 (same attribute names
and shapes).  The tables are computed by the native library in extended
precision instead of being read from precomputed mpmath files.
"""
import ctypes as C
import functools

import numpy as np

from . import _capi as A


@functools.lru_cache(maxsize=None)
def _ref_tables(N):
    L = A.lib()
    Np = (N + 1) * (N + 2) * (N + 3) // 6
    Nfp = (N + 1) * (N + 2) // 2
    shapes = dict(xyz=(Np, 3), rst=(Np, 3), V=(Np, Np), M=(Np, Np), invM=(Np, Np), Dr=(Np, Np), Ds=(Np, Np),
                  Dt=(Np, Np), WDr=(Np, Np), WDs=(Np, Np), WDt=(Np, Np), Emat=(Np, 4 * Nfp), LIFT=(Np, 4 * Nfp))
    t = {k: np.zeros(s) for k, s in shapes.items()}
    t["Fmask"] = np.zeros((4, Nfp), np.int32)
    t["FmaskP"] = np.zeros((4, 3, Nfp), np.int32)
    rt = A.RefTables()
    for k in shapes:
        setattr(rt, k, A.dptr(t[k]))
    rt.Fmask = A.iptr(t["Fmask"])
    rt.FmaskP = A.iptr(t["FmaskP"])
    A.check(L.dglm_ref_operators(N, C.byref(rt)))
    for v in t.values():
        v.setflags(write=False)
    return t


@functools.lru_cache(maxsize=None)
def _interp_tables(N, high_N):
    Np = (N + 1) * (N + 2) * (N + 3) // 6
    hNp = (high_N + 1) * (high_N + 2) * (high_N + 3) // 6
    hr = np.zeros((hNp, 3))
    iM = np.zeros((hNp, Np))
    hM = np.zeros((hNp, hNp))
    A.check(A.lib().dglm_interp_operators(N, high_N, A.dptr(hr), A.dptr(iM), A.dptr(hM)))
    for v in (hr, iM, hM):
        v.setflags(write=False)
    return hr, iM, hM


def gll_nodes(N):
    x = np.zeros(N + 1)
    A.check(A.lib().dglm_gll_nodes(N, A.dptr(x)))
    return x


def jacobi(N, a, b, x):
    """Orthonormal Jacobi polynomial and its derivative at points x (jacobi.py:69-114)."""
    x = np.ascontiguousarray(x, dtype=np.float64)
    P = np.zeros_like(x)
    dP = np.zeros_like(x)
    A.check(A.lib().dglm_jacobi(N, C.c_double(a), C.c_double(b), x.size, A.dptr(x), A.dptr(P), A.dptr(dP)))
    return P, dP


class TetraRef:
    """Reference tetrahedron of order N (tetra.py:305-331)."""

    def __init__(self, N, interpN=2, use_mp=False):
        self.N = N
        self.Np = (N + 1) * (N + 2) * (N + 3) // 6
        self.Nface = 4
        self.Nfp = (N + 1) * (N + 2) // 2
        self.Nrot = 3
        self.interpN = interpN
        self.high_N = N + interpN
        self.high_Np = (self.high_N + 1) * (self.high_N + 2) * (self.high_N + 3) // 6
        self.use_mp = use_mp
        t = _ref_tables(N)
        self.xyz = t["xyz"]
        self.rst = t["rst"]
        self.high_rst = _interp_tables(N, self.high_N)[0]


class Operator3DRef:
    """Volume operators (operator_3d_ref.py:302-349)."""

    def __init__(self, tet):
        self.tet = tet
        self.N = tet.N
        self.Np = tet.Np
        self.high_N = tet.high_N
        self.high_Np = tet.high_Np
        t = _ref_tables(tet.N)
        self.V = t["V"]
        self.M = t["M"]
        self.invM = t["invM"]
        self.Dr, self.Ds, self.Dt = t["Dr"], t["Ds"], t["Dt"]
        self.WDr, self.WDs, self.WDt = t["WDr"], t["WDs"], t["WDt"]
        _, self.interpM, self.high_M = _interp_tables(tet.N, tet.high_N)


class Operator3DFace:
    """Face masks and lift (operator_3d_face.py:275-301)."""

    def __init__(self, op_ref):
        self.op_ref = op_ref
        self.tet = op_ref.tet
        t = _ref_tables(self.tet.N)
        self.Fmask = t["Fmask"]
        self.FmaskP = t["FmaskP"]
        self.Emat = t["Emat"]
        self.LIFT = t["LIFT"]
