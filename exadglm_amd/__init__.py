"""exadglm_amd: MI355X-native explicit DGLM solver core (advection, compressible Euler).

The solver runs on the GPU through libexadglm_hip.so (HIP kernels for gfx950
behind the C ABI of include/exadglm.h).  This package holds the Python
binding and the host-side mirror of the reference's driver API.
"""
from . import _capi  # noqa: F401
from .problem import Problem, cube_problem, uniform_problem  # noqa: F401
from .reference import Operator3DFace, Operator3DRef, TetraRef  # noqa: F401
from .mesh import Mesh3D, Uniform3DSingle  # noqa: F401
from .solver import Solver  # noqa: F401

__version__ = "0.1.0"
