"""Reference-element tables: known answers from the reference's own tests,
closed forms, and (where the reference checkout exists) its shipped tables."""
import json
import os

import numpy as np
import pytest
from numpy.testing import assert_array_almost_equal as aae
from numpy.testing import assert_array_equal as ae

from conftest import REFERENCE, has_reference
from exadglm_amd.reference import Operator3DFace, Operator3DRef, TetraRef, gll_nodes, jacobi

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def ops(N):
    tet = TetraRef(N)
    o = Operator3DRef(tet)
    return tet, o, Operator3DFace(o)


# Known answers of dim3/tests/test_jacobi.py:10-51.
def test_gll_known_answers():
    ae(gll_nodes(1), [-1, 1])
    ae(gll_nodes(2), [-1, 0, 1])
    aae(gll_nodes(3), [-1, -0.447213595499958, 0.4472135954999580, 1], 15)
    aae(gll_nodes(4), [-1, -0.654653670707977, 0, 0.654653670707977, 1], 15)
    aae(gll_nodes(5), [-1, -0.765055323929465, -0.285231516480645, 0.285231516480645, 0.765055323929465, 1], 15)


@pytest.mark.parametrize("N,P,dP,dec", [
    (2, [1.581138830084190, -0.790569415042095, 1.581138830084190], [-4.743416490252569, 0, 4.743416490252569], 15),
    (3, [-1.87082869338697, 0.836660026534076, -0.836660026534076, 1.87082869338697],
     [11.224972160321824, 0, 0, 11.224972160321824], 15),
    (4, [2.121320343559642, -0.909137290096990, 0.795495128834866, -0.909137290096990, 2.121320343559642],
     [-21.213203435596430, 0, 0, 0, 21.213203435596430], 14),
    (5, [-2.345207879911714, 0.984276557099483, -0.812914072195837, 0.812914072195837, -0.984276557099483,
         2.345207879911714], [35.178118198675726, 0, 0, 0, 0, 35.178118198675726], 14),
])
def test_jacobi_known_answers(N, P, dP, dec):
    p, dp = jacobi(N, 0, 0, gll_nodes(N))
    aae(p, P, dec)
    aae(dp, dP, dec)


# Known answers of dim3/tests/test_operator_3d_face.py:37-64.
def test_face_masks_N3():
    _, _, f = ops(3)
    ae(f.Fmask[0], [3, 2, 1, 0, 6, 5, 4, 8, 7, 9])
    ae(f.FmaskP[0, 0], [0, 1, 2, 3, 4, 5, 6, 7, 8, 9])
    ae(f.FmaskP[0, 1], [3, 6, 8, 9, 2, 5, 7, 1, 4, 0])
    ae(f.FmaskP[0, 2], [9, 7, 4, 0, 8, 5, 1, 6, 2, 3])
    ae(f.Fmask[1], [0, 1, 2, 3, 10, 11, 12, 16, 17, 19])
    ae(f.FmaskP[1, 0], [3, 2, 1, 0, 12, 11, 10, 17, 16, 19])
    ae(f.FmaskP[1, 1], [0, 10, 16, 19, 1, 11, 17, 2, 12, 3])
    ae(f.FmaskP[1, 2], [19, 17, 12, 3, 16, 11, 2, 10, 1, 0])
    ae(f.Fmask[2], [3, 6, 8, 9, 12, 14, 15, 17, 18, 19])
    ae(f.FmaskP[2, 0], [9, 8, 6, 3, 15, 14, 12, 18, 17, 19])
    ae(f.FmaskP[2, 1], [3, 12, 17, 19, 6, 14, 18, 8, 15, 9])
    ae(f.FmaskP[2, 2], [19, 18, 15, 9, 17, 14, 8, 12, 6, 3])
    ae(f.Fmask[3], [9, 7, 4, 0, 15, 13, 10, 18, 16, 19])
    ae(f.FmaskP[3, 0], [0, 4, 7, 9, 10, 13, 15, 16, 18, 19])
    ae(f.FmaskP[3, 1], [9, 15, 18, 19, 7, 13, 16, 4, 10, 0])
    ae(f.FmaskP[3, 2], [19, 16, 10, 0, 18, 13, 4, 15, 7, 9])


@pytest.mark.parametrize("N", [1, 2, 3, 4, 5, 6])
def test_closed_forms(N):
    tet, o, f = ops(N)
    Np = tet.Np
    # volume of the reference tet is 4/3; face areas 2 each (test_operator_3d_ref.py:40-47, _face.py:67-76)
    assert abs(o.M.sum() - 4 / 3) < 1e-13
    assert abs(f.Emat.sum() - 8) < 1e-12
    # weak operators are M^-1 D^T M and LIFT = M^-1 Emat
    Mi = np.linalg.inv(o.M)
    for D, W in ((o.Dr, o.WDr), (o.Ds, o.WDs), (o.Dt, o.WDt)):
        assert np.abs(Mi @ D.T @ o.M - W).max() < 1e-11 * max(1, np.abs(W).max())
        assert np.abs(D.sum(axis=1)).max() < 1e-11  # derivative of a constant
    assert np.abs(Mi @ f.Emat - f.LIFT).max() < 1e-11 * np.abs(f.LIFT).max()
    # derivative exactness on monomials of degree <= N
    r, s, t = tet.rst.T
    for (a, b, c) in [(N, 0, 0), (0, N, 0), (0, 0, N), (N - 1, 1, 0) if N > 1 else (1, 0, 0)]:
        p = r ** a * s ** b * t ** c
        dr = a * r ** max(a - 1, 0) * s ** b * t ** c
        ds = b * r ** a * s ** max(b - 1, 0) * t ** c
        dt = c * r ** a * s ** b * t ** max(c - 1, 0)
        for D, ex in ((o.Dr, dr), (o.Ds, ds), (o.Dt, dt)):
            assert np.abs(D @ p - ex).max() < 1e-11
    # Emat face blocks are symmetric positive definite only in Fmask order
    for fc in range(4):
        B = f.Emat[f.Fmask[fc], fc * tet.Nfp:(fc + 1) * tet.Nfp]
        assert np.allclose(B, B.T, atol=1e-14)
        assert np.linalg.eigvalsh(B).min() > 0
    # interpolation reproduces the high-order node coordinates
    hr = tet.high_rst
    assert np.abs(o.interpM @ tet.rst - hr).max() < 1e-12
    assert abs(o.high_M.sum() - 4 / 3) < 1e-12
    assert Np == (N + 1) * (N + 2) * (N + 3) // 6


def _checksum(a):
    a = np.asarray(a, dtype=np.float64).ravel()
    w = np.cos(np.arange(a.size) * 0.7311)
    return [float(np.sum(a * a)), float(np.sum(a * w))]


@pytest.mark.parametrize("N", list(range(1, 11)))
def test_against_reference_table_checksums(N):
    """Checksums of the reference's shipped dim3/data/bin_mats tables (fixture
    written by tests/golden/make_operator_fixture.py)."""
    with open(os.path.join(GOLDEN, "operator_checksums.json")) as fh:
        gold = json.load(fh)[str(N)]
    tet, o, f = ops(N)
    mine = dict(rst=tet.rst, M=o.M, Dr=o.Dr, Ds=o.Ds, Dt=o.Dt, WDr=o.WDr, WDs=o.WDs, WDt=o.WDt, Emat=f.Emat,
                LIFT=f.LIFT, interpM2=o.interpM)
    for name, (ss, sw) in gold.items():
        c = _checksum(mine[name])
        assert abs(c[0] - ss) <= 1e-13 * max(1, abs(ss)), name
        assert abs(c[1] - sw) <= 1e-12 * max(1, np.sqrt(abs(ss))), name


@pytest.mark.skipif(not has_reference(), reason="reference checkout not present")
@pytest.mark.parametrize("N", list(range(1, 11)))
def test_against_reference_tables(N):
    tet, o, f = ops(N)
    Np, Nfp = tet.Np, tet.Nfp
    d = os.path.join(REFERENCE, "dim3/data/bin_mats", f"N{N}")
    # the Vandermonde conditioning grows with N: 3.2e-15 and 3.7e-15 at N = 9, 10
    tol = 3e-15 if N <= 8 else 5e-15
    for name, mine in dict(rst=tet.rst, M=o.M, Dr=o.Dr, WDr=o.WDr, WDs=o.WDs, WDt=o.WDt, Emat=f.Emat,
                           LIFT=f.LIFT).items():
        ref = np.fromfile(os.path.join(d, f"{name}.bin")).reshape(mine.shape)
        assert np.abs(ref - mine).max() <= tol * np.abs(ref).max(), name
