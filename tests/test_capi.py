"""The C ABI library loads, exports every declared symbol, and rejects bad
input with error codes (no GPU needed: these fail before any device call)."""
import ctypes as C
import math
import os
import re
import tempfile

import numpy as np
import pytest

import exadglm_amd._capi as A
from exadglm_amd.problem import cube_problem
import pyoracle as O

HEADER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include", "exadglm.h")


def declared():
    src = open(HEADER).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(dglm_\w+)\s*\(", src)))


def test_every_declared_symbol_is_exported():
    L = A.lib()
    names = declared()
    assert len(names) >= 30
    for n in names:
        assert hasattr(L, n), n
    assert sorted(A.EXPORTS) == names


def params(**kw):
    p = A.Params()
    p.N, p.physics, p.precision, p.flux = 3, A.DGLM_EULER, A.DGLM_FP64, A.DGLM_FLUX_LM
    p.cfl, p.gamma, p.tau_scale, p.tmax, p.device = 0.5, 1.4, 1.2, math.inf, 0
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def test_create_rejects_bad_params():
    L = A.lib()
    ctx = C.c_void_p()
    m = A.Mesh()
    m.K = 6
    for kw, code in [(dict(physics=7), A.DGLM_ERR_ARG), (dict(precision=9), A.DGLM_ERR_ARG), (dict(precision=4), A.DGLM_ERR_ARG),
                     (dict(N=11), A.DGLM_ERR_UNSUPPORTED), (dict(N=7, precision=A.DGLM_FP32), A.DGLM_ERR_UNSUPPORTED),
                     (dict(N=7, precision=A.DGLM_BF16X3), A.DGLM_ERR_UNSUPPORTED),
                     (dict(N=7, precision=A.DGLM_BF16X6), A.DGLM_ERR_UNSUPPORTED),
                     (dict(cfl=-1.0), A.DGLM_ERR_ARG),
                     (dict(gamma=1.0), A.DGLM_ERR_ARG)]:
        rc = L.dglm_create(C.byref(params(**kw)), C.byref(m), C.byref(ctx))
        assert rc == code, (kw, rc, L.dglm_last_error())
        assert L.dglm_last_error()
    # missing arrays
    assert L.dglm_create(C.byref(params()), C.byref(m), C.byref(ctx)) == A.DGLM_ERR_ARG


def test_create_from_dir_reports_missing_and_truncated_files():
    L = A.lib()
    ctx = C.c_void_p()
    with tempfile.TemporaryDirectory() as d:
        rc = L.dglm_create_from_dir(C.byref(params()), d.encode(), 0, C.byref(ctx))
        assert rc == A.DGLM_ERR_IO and b"dim.bin" in L.dglm_last_error()
        pb = cube_problem(3, 2)
        O.write_datain(pb, d)
        with open(os.path.join(d, "vmapP.bin"), "r+b") as f:
            f.truncate(100)
        rc = L.dglm_create_from_dir(C.byref(params()), d.encode(), 0, C.byref(ctx))
        assert rc == A.DGLM_ERR_IO and b"vmapP.bin" in L.dglm_last_error()


def test_create_rejects_inconsistent_connectivity():
    """A vmapP that is not a face rotation of the neighbour is refused."""
    L = A.lib()
    from exadglm_amd.solver import make_mesh_struct
    pb = cube_problem(3, 2)
    arr = dict(pb.arrays)
    arr["vmapP"] = np.ascontiguousarray(np.roll(pb.arrays["vmapP"], 1, axis=2))
    m = make_mesh_struct(arr, pb.K)
    ctx = C.c_void_p()
    rc = L.dglm_create(C.byref(params()), C.byref(m), C.byref(ctx))
    assert rc == A.DGLM_ERR_UNSUPPORTED and b"rotation" in L.dglm_last_error()


def test_null_arguments():
    L = A.lib()
    assert L.dglm_step(None, 1, None) == A.DGLM_ERR_ARG
    assert L.dglm_get_state(None, None) == A.DGLM_ERR_ARG
    assert L.dglm_ref_operators(3, None) == A.DGLM_ERR_ARG
    assert L.dglm_uniform_mesh(1, 2, 2, None, None, None, None, None, None, None, None) == A.DGLM_ERR_ARG


REALS = ("WDr", "WDs", "WDt", "LIFT", "Fscale", "rx", "ry", "rz", "sx", "sy", "sz", "tx", "ty", "tz", "nx", "ny",
         "nz", "px", "py", "pz")


def test_create_from_dir_reads_f4_and_refuses_mixed_widths():
    """datain written with dtype 'f4' (REAL = float) is read; one f8 file
    among f4 ones, or a file of neither size, is DGLM_ERR_IO naming it.
    (Without a GPU the accepted directory stops at the first device call.)"""
    L = A.lib()
    ctx = C.c_void_p()
    pb = cube_problem(3, 2)
    with tempfile.TemporaryDirectory() as d:
        O.write_datain(pb, d)
        for n in REALS:
            np.fromfile(os.path.join(d, n + ".bin")).astype(np.float32).tofile(os.path.join(d, n + ".bin"))
        rc = L.dglm_create_from_dir(C.byref(params()), d.encode(), 0, C.byref(ctx))
        assert rc != A.DGLM_ERR_IO, L.dglm_last_error()
        if rc == A.DGLM_OK:
            L.dglm_destroy(ctx)
        np.asarray(pb.arrays["ry"], dtype=np.float64).tofile(os.path.join(d, "ry.bin"))
        rc = L.dglm_create_from_dir(C.byref(params()), d.encode(), 0, C.byref(ctx))
        assert rc == A.DGLM_ERR_IO and b"ry.bin" in L.dglm_last_error() and b"f8" in L.dglm_last_error()
        np.zeros(pb.K * 2 + 1, np.float32).tofile(os.path.join(d, "ry.bin"))
        rc = L.dglm_create_from_dir(C.byref(params()), d.encode(), 0, C.byref(ctx))
        assert rc == A.DGLM_ERR_IO and b"ry.bin" in L.dglm_last_error()
