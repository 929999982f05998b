"""ctypes binding of libexadglm_hip.so (the C ABI declared in include/exadglm.h).

The library is loaded from this package directory only; a missing library is an
ImportError-style failure (there is no fallback implementation).
"""
import ctypes as C
import os

import numpy as np

_PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_PKG, "libexadglm_hip.so")

DGLM_OK = 0
DGLM_ERR_ARG = -1
DGLM_ERR_HIP = -2
DGLM_ERR_IO = -3
DGLM_ERR_STATE = -4
DGLM_ERR_UNSUPPORTED = -5

DGLM_ADVECTION = 0
DGLM_EULER = 1
DGLM_FP64 = 0
DGLM_FP32 = 1
DGLM_BF16X3 = 2
DGLM_BF16X6 = 3
DGLM_FLUX_LM = 0
DGLM_FLUX_RUSANOV = 1
DGLM_STAGE_AUTO = 0
DGLM_STAGE_SIMPLE = 1
DGLM_STAGE_MFMA = 2
DGLM_STAGE_MFMA_WS = 3
DGLM_STAGE_MFMA_MX = 4

_ERRNAMES = {
    DGLM_ERR_ARG: "DGLM_ERR_ARG",
    DGLM_ERR_HIP: "DGLM_ERR_HIP",
    DGLM_ERR_IO: "DGLM_ERR_IO",
    DGLM_ERR_STATE: "DGLM_ERR_STATE",
    DGLM_ERR_UNSUPPORTED: "DGLM_ERR_UNSUPPORTED",
}


class DGLMError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{_ERRNAMES.get(code, code)}: {msg}")
        self.code = code


PD = C.POINTER(C.c_double)
PI = C.POINTER(C.c_int)


class RefTables(C.Structure):
    _fields_ = [(n, PD) for n in ("xyz", "rst", "V", "M", "invM", "Dr", "Ds", "Dt", "WDr", "WDs", "WDt",
                                  "Emat", "LIFT")] + [("Fmask", PI), ("FmaskP", PI)]


class Params(C.Structure):
    _fields_ = [("N", C.c_int), ("physics", C.c_int), ("precision", C.c_int), ("flux", C.c_int),
                ("cfl", C.c_double), ("gamma", C.c_double), ("tau_scale", C.c_double), ("tmax", C.c_double),
                ("device", C.c_int)]


class Mesh(C.Structure):
    _fields_ = ([("K", C.c_int), ("matrix_order", C.c_int)]
                + [(n, PI) for n in ("EtoB", "vmapP", "vmapF", "Fmask")]
                + [(n, PD) for n in ("WDr", "WDs", "WDt", "LIFT", "Fscale", "rx", "ry", "rz", "sx", "sy", "sz",
                                     "tx", "ty", "tz", "nx", "ny", "nz", "px", "py", "pz")])


# Names of every function the header declares (checked by the tests).
EXPORTS = [
    "dglm_last_error", "dglm_version", "dglm_ref_sizes", "dglm_ref_operators", "dglm_interp_operators",
    "dglm_gll_nodes", "dglm_jacobi", "dglm_uniform_mesh", "dglm_mesh_points", "dglm_mesh_metrics",
    "dglm_mesh_normals", "dglm_mesh_vmaps", "dglm_mesh_spatial_h", "dglm_create", "dglm_create_from_dir",
    "dglm_destroy", "dglm_sizes", "dglm_set_state", "dglm_get_state", "dglm_set_state_device",
    "dglm_get_state_device", "dglm_set_velocity", "dglm_set_time", "dglm_get_time", "dglm_rhs", "dglm_step",
    "dglm_integrals", "dglm_sync", "dglm_set_timing", "dglm_get_timing",
    "dglm_set_stage_kernel", "dglm_get_stage_kernel", "dglm_adjoint_rhs", "dglm_adjoint_record",
    "dglm_adjoint_backward", "dglm_adjoint_set_exact", "dglm_adjoint_set_keep", "dglm_adjoint_tape_bytes",
    "dglm_set_graphs",
    "dglm_dirichlet_faces", "dglm_set_dirichlet", "dglm_set_dirichlet_device", "dglm_set_alloc_faults",
    "dglm_step_noalloc",
]

# Limits of the allocation-free entry point (as the header's #defines).
DGLM_NOALLOC_MAX_STEPS = 65536
DGLM_NOALLOC_MAX_BLOCKS = 262144

_lib = None


def lib():
    """Load the in-tree library (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: run __graft_entry__.build() (or python -m exadglm_amd.build)")
        L = C.CDLL(LIB_PATH)
        L.dglm_last_error.restype = C.c_char_p
        _lib = L
    return _lib


def check(rc):
    if rc != DGLM_OK:
        raise DGLMError(rc, lib().dglm_last_error().decode())
    return rc


def dptr(a):
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags.c_contiguous, "expected C-contiguous float64 array"
    return a.ctypes.data_as(PD)


def iptr(a):
    if a is None:
        return None
    assert a.dtype == np.int32 and a.flags.c_contiguous, "expected C-contiguous int32 array"
    return a.ctypes.data_as(PI)
