"""Uniform periodic tetrahedral mesh with the reference's Python API.

Uniform3DSingle mirrors dim3/src/python/uniform_3d_single.py:8-134 and Mesh3D
the serial path of dim3/src/python/mesh_3d.py:9-116 (same attribute names and
shapes); the work is done by the native library.
"""
import re

import numpy as np

from . import _capi as A


def _phy_tag(spec):
    if spec == "pbc":
        return -1
    m = re.findall(r"bc(\d+)", spec)
    if not m:
        raise ValueError(f"boundary spec must be 'pbc' or 'bc<n>', got {spec!r}")
    return int(m[0])


class Uniform3DSingle:
    """(nx-1)(ny-1)(nz-1) hexes split into 6 tetrahedra each."""

    def __init__(self, nx, ny, nz, x1, x2, y1, y2, z1, z2, bc=None, mod_suffix="cpp"):
        bc = bc or {"x": "pbc", "y": "pbc", "z": "pbc"}
        self.nx, self.ny, self.nz = nx, ny, nz
        self.x1, self.x2, self.y1, self.y2, self.z1, self.z2 = x1, x2, y1, y2, z1, z2
        self.bc = bc
        self.Ncube = (nx - 1) * (ny - 1) * (nz - 1)
        self.Nelem = self.Ncube * 6
        self.Nface = 4
        K = self.Nelem
        self.VXYZ = np.zeros((nx * ny * nz, 3))
        self.EtoV = np.zeros((K, 4), np.int32)
        self.EtoE = np.zeros((K, 4), np.int32)
        self.EtoF = np.zeros((K, 4), np.int32)
        self.EtoR = np.zeros((K, 4), np.int32)
        self.EtoB = np.zeros((K, 4), np.int32)
        box = np.array([x1, x2, y1, y2, z1, z2], dtype=np.float64)
        tags = np.array([_phy_tag(bc["x"]), _phy_tag(bc["y"]), _phy_tag(bc["z"])], np.int32)
        A.check(A.lib().dglm_uniform_mesh(nx, ny, nz, A.dptr(box), A.iptr(tags), A.dptr(self.VXYZ),
                                          A.iptr(self.EtoV), A.iptr(self.EtoE), A.iptr(self.EtoF),
                                          A.iptr(self.EtoR), A.iptr(self.EtoB)))


class Mesh3D:
    """Node coordinates, affine metrics, normals and vmaps of a single partition."""

    def __init__(self, op_face, mesh_elem, mod_suffix="cpp"):
        L = A.lib()
        self.op_face = op_face
        self.op_ref = op_face.op_ref
        self.tet = op_face.tet
        self.mesh_elem = mesh_elem
        self.Nelem = K = mesh_elem.Nelem
        self.Nface = 4
        self.Np = Np = self.tet.Np
        self.Nfp = Nfp = self.tet.Nfp
        self.VX = np.ascontiguousarray(mesh_elem.VXYZ[:, 0])
        self.VY = np.ascontiguousarray(mesh_elem.VXYZ[:, 1])
        self.VZ = np.ascontiguousarray(mesh_elem.VXYZ[:, 2])
        self.EtoV, self.EtoE, self.EtoF = mesh_elem.EtoV, mesh_elem.EtoE, mesh_elem.EtoF
        self.EtoR, self.EtoB = mesh_elem.EtoR, mesh_elem.EtoB
        self._VXYZ = np.ascontiguousarray(mesh_elem.VXYZ, dtype=np.float64)

        self.PX, self.PY, self.PZ = self.make_inner_points(self.tet.rst)
        self.J = np.zeros(K)
        metrics = np.zeros((9, K))
        A.check(L.dglm_mesh_metrics(K, A.iptr(self.EtoV), A.dptr(self._VXYZ), A.dptr(self.J), A.dptr(metrics)))
        (self.rx, self.ry, self.rz, self.sx, self.sy, self.sz, self.tx, self.ty, self.tz) = metrics
        self._metrics = metrics
        self.NX, self.NY, self.NZ, self.sJ, self.Fscale = (np.zeros((K, 4)) for _ in range(5))
        A.check(L.dglm_mesh_normals(K, A.dptr(self.J), A.dptr(metrics), A.dptr(self.NX), A.dptr(self.NY),
                                    A.dptr(self.NZ), A.dptr(self.sJ), A.dptr(self.Fscale)))
        N = self.tet.N
        self.dt_scale = 1 / (np.max(self.Fscale) * N * N)
        self.myrank = -1
        self.vmapM = np.zeros((K, 4, Nfp), np.int32)
        self.vmapP = np.zeros((K, 4, Nfp), np.int32)
        self.vmapF = np.zeros((K, 4), np.int32)
        A.check(L.dglm_mesh_vmaps(K, N, A.iptr(self.EtoE), A.iptr(self.EtoF), A.iptr(self.EtoR),
                                  A.iptr(self.vmapM), A.iptr(self.vmapP), A.iptr(self.vmapF)))
        self.high_PX, self.high_PY, self.high_PZ = self.make_inner_points(self.tet.high_rst)

    def make_inner_points(self, rst):
        rst = np.ascontiguousarray(rst, dtype=np.float64)
        K, Np = self.Nelem, rst.shape[0]
        PX, PY, PZ = (np.zeros((K, Np)) for _ in range(3))
        A.check(A.lib().dglm_mesh_points(K, Np, A.iptr(self.EtoV), A.dptr(self._VXYZ), A.dptr(rst), A.dptr(PX),
                                         A.dptr(PY), A.dptr(PZ)))
        return PX, PY, PZ

    def calc_spatial_h(self, verbose=0):
        h = np.zeros(self.Nelem)
        A.check(A.lib().dglm_mesh_spatial_h(self.Nelem, A.iptr(self.EtoV), A.dptr(self._VXYZ), A.dptr(h)))
        if verbose >= 1:
            hJ = self.J / self.sJ.min(axis=1)
            print(f"min(h) ={h.min():1.5f}, max(h) ={h.max():1.5f}")
            print(f"min(hJ)={hJ.min():1.5f}, max(hJ)={hJ.max():1.5f}")
        return h.min(), h.max()
