"""Uniform mesh and geometry: known answers of dim3/tests/test_uniform_3d_single.py,
invariants of test_mesh_3d_common.py / test_mesh_3d_serial_uniform_pbc.py."""
import numpy as np
import pytest
from numpy.testing import assert_array_almost_equal as aae
from numpy.testing import assert_array_equal as ae

from exadglm_amd.mesh import Mesh3D, Uniform3DSingle
from exadglm_amd.reference import Operator3DFace, Operator3DRef, TetraRef
import exadglm_amd._capi as A

# the reference test mesh: x in [0,2] nx=3, y in [1,3] ny=3, z in [2,5] nz=4
MSH = dict(nx=3, ny=3, nz=4, x1=0, x2=2, y1=1, y2=3, z1=2, z2=5)


def small(bc=None):
    return Uniform3DSingle(MSH["nx"], MSH["ny"], MSH["nz"], MSH["x1"], MSH["x2"], MSH["y1"], MSH["y2"], MSH["z1"],
                           MSH["z2"], bc)


def test_vxyz_known_answers():
    m = small()
    for i, v in {0: [0, 1, 2], 1: [0, 1, 3], 2: [0, 1, 4], 3: [0, 1, 5], 4: [0, 2, 2], 5: [0, 2, 3], 6: [0, 2, 4],
                 7: [0, 2, 5], 8: [0, 3, 2], 9: [0, 3, 3], 22: [1, 3, 4], 23: [1, 3, 5]}.items():
        ae(m.VXYZ[i], v)


def test_etov_known_answers():
    m = small()
    exp = {0: [0, 17, 5, 1], 1: [0, 17, 4, 5], 2: [0, 17, 16, 4], 3: [0, 17, 12, 16], 4: [0, 17, 13, 12],
           5: [0, 17, 1, 13], 6: [1, 18, 6, 2], 7: [1, 18, 5, 6], 8: [1, 18, 17, 5], 9: [1, 18, 13, 17],
           10: [1, 18, 14, 13], 11: [1, 18, 2, 14], 30: [6, 23, 11, 7], 31: [6, 23, 10, 11], 32: [6, 23, 22, 10],
           33: [6, 23, 18, 22], 34: [6, 23, 19, 18], 35: [6, 23, 7, 19], 48: [14, 31, 19, 15], 49: [14, 31, 18, 19],
           50: [14, 31, 30, 18], 51: [14, 31, 26, 30], 52: [14, 31, 27, 26], 53: [14, 31, 15, 27],
           66: [18, 35, 23, 19], 67: [18, 35, 22, 23], 68: [18, 35, 34, 22], 69: [18, 35, 30, 34],
           70: [18, 35, 31, 30], 71: [18, 35, 19, 31]}
    for e, v in exp.items():
        ae(m.EtoV[e], v)


@pytest.mark.parametrize("bc", [None, {"x": "bc1", "y": "pbc", "z": "pbc"}, {"x": "pbc", "y": "bc2", "z": "bc3"}])
def test_neighbour_involution(bc):
    m = small(bc)
    for e in range(m.Nelem):
        for f in range(4):
            nb = m.EtoE[e, f]
            if nb < 0:
                assert m.EtoB[e, f] >= 0
                continue
            nf, r = m.EtoF[e, f], m.EtoR[e, f]
            assert m.EtoE[nb, nf] == e and m.EtoF[nb, nf] == f and m.EtoR[nb, nf] == r
    if bc is None:
        assert (m.EtoB == -99).all()


def geometry(N=3, nx=4, ny=5, nz=6, box=(0.0, 1.0, 0.0, 1.2, 0.0, 1.4), bc=None):
    tet = TetraRef(N)
    of = Operator3DFace(Operator3DRef(tet))
    return Mesh3D(of, Uniform3DSingle(nx, ny, nz, *box, bc=bc))


def test_jacobian_sum_and_surface():
    m = geometry()
    # box volume 1.68 over reference volume 4/3 (test_mesh_3d_common.py:41-47)
    aae(np.sum(m.J), 1.26, 13)
    faces = [(1, 0, 2), (0, 1, 3), (1, 2, 3), (2, 0, 3)]
    for e in range(0, m.Nelem, 7):
        vx = np.column_stack([m.VX[m.EtoV[e]], m.VY[m.EtoV[e]], m.VZ[m.EtoV[e]]])
        for f, (a, b, c) in enumerate(faces):
            v1, v2 = vx[b] - vx[a], vx[c] - vx[a]
            area = 0.5 * np.linalg.norm(np.cross(v1, v2))
            aae(m.sJ[e, f], area / 2, 14)
            n = np.array([m.NX[e, f], m.NY[e, f], m.NZ[e, f]])
            aae(np.dot(n, v1 + v2), 0, 14)
            assert np.dot(n, np.cross(v1, v2)) > 0
            aae(np.linalg.norm(n), 1, 14)
    aae(m.Fscale, m.sJ / m.J[:, None], 14)


def test_vmaps_match_modulo_period():
    m = geometry()
    L = np.array([1.0, 1.2, 1.4])
    X = np.stack([m.PX.ravel(), m.PY.ravel(), m.PZ.ravel()], 1)
    d = X[m.vmapM.ravel()] - X[m.vmapP.ravel()]
    d -= L * np.round(d / L)
    assert np.abs(d).max() < 1e-12
    # vmapF points at the partner face
    for e in range(m.Nelem):
        for f in range(4):
            ae(m.vmapF[e, f], m.EtoE[e, f] * 4 + m.EtoF[e, f])


def test_high_points_interpolate():
    m = geometry(N=2)
    aae(m.high_PX.T, m.op_ref.interpM @ m.PX.T, 14)


def test_spatial_h_inscribed_sphere():
    # dim3/tests/test_mesh_3d_spatial_h.py: faces x=0,y=0,z=0, 2x+6y+3z=14 -> diameter 14/9
    V = np.array([[0, 0, 0], [7, 0, 0], [0, 7 / 3, 0], [0, 0, 14 / 3]], dtype=np.float64)
    E = np.array([[0, 1, 2, 3]], np.int32)
    h = np.zeros(1)
    A.check(A.lib().dglm_mesh_spatial_h(1, A.iptr(E), A.dptr(V), A.dptr(h)))
    aae(h[0], 2 * 7 / 9, 15)
