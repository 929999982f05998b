"""The allocation-free entry point dglm_step_noalloc against dglm_step: the
same kernels on the same inputs, so states and dt histories are bitwise
equal; two calls in a row on one device, the second on a smaller mesh and
fewer steps (the shared scratch array still holds the first call's larger
contents), and with every library allocation made to fail meanwhile (the
entry point makes none)."""
import numpy as np
import pytest

import exadglm_amd._capi as A
from helpers import smooth_state
from exadglm_amd.problem import cube_problem
from exadglm_amd.solver import Solver

pytestmark = pytest.mark.gpu


def _run(pb, physics, u0, nsteps, noalloc, hist):
    with Solver(pb, physics, cfl=0.5) as s:
        s.set_state(u0)
        if noalloc:
            A.check(A.lib().dglm_set_alloc_faults(1000))
            try:
                d = s.step_noalloc(nsteps, return_dt=hist)
            finally:
                A.check(A.lib().dglm_set_alloc_faults(0))
        else:
            d = s.step(nsteps, return_dt=hist)
        return s.get_state(), d


@pytest.mark.parametrize("physics", ["euler", "advection"])
def test_step_noalloc_equals_step(gpu, physics):
    cases = [(4, 6, 12, 11), (3, 3, 5, 12)]  # (N, hexes per side, steps, seed): the second call is smaller
    for N, nhex, nsteps, seed in cases:
        print(f"noalloc {physics} N={N} {nhex}^3 hexes {nsteps} steps", flush=True)
        pb = cube_problem(N, nhex)
        u0 = smooth_state(pb, physics, seed=seed)
        for hist in (True, False):
            u_ref, d_ref = _run(pb, physics, u0, nsteps, False, hist)
            u_na, d_na = _run(pb, physics, u0, nsteps, True, hist)
            np.testing.assert_array_equal(u_na, u_ref)
            if hist:
                np.testing.assert_array_equal(d_na, d_ref)
                assert (d_na > 0).all()


def test_step_noalloc_limits(gpu):
    """Sizes beyond the scratch array are refused before any device work."""
    pb = cube_problem(3, 2)
    u0 = smooth_state(pb, "euler", seed=3)
    with Solver(pb, "euler") as s:
        s.set_state(u0)
        with pytest.raises(A.DGLMError) as ei:
            s.step_noalloc(A.DGLM_NOALLOC_MAX_STEPS + 1)
        assert ei.value.code == A.DGLM_ERR_ARG and "DGLM_NOALLOC_MAX_STEPS" in str(ei.value)
        np.testing.assert_array_equal(s.get_state(), u0)  # nothing ran
