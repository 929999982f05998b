"""The host-side C++ (operator tables, uniform mesh, geometry, vmaps, the
datain reader and the connectivity packing) under AddressSanitizer and UBSan:
a self-check program built from the same sources with g++
-fsanitize=address,undefined (the reference's build(use_sanitize=True) does
the same for its CPU solver, dim3/src/python/dglm_3d_launcher.py:389-390),
run on generated meshes, on datain directories in both dtypes, and on
corrupted ones, which must give a clean error (exit status 2), never a
sanitizer report."""
import os
import tempfile

import numpy as np
import pytest

from exadglm_amd import build as B
from exadglm_amd.problem import cube_problem, uniform_problem
import pyoracle as O

REALS = ("WDr", "WDs", "WDt", "LIFT", "Fscale", "rx", "ry", "rz", "sx", "sy", "sz", "tx", "ty", "tz", "nx", "ny",
         "nz", "px", "py", "pz")


@pytest.fixture(scope="module")
def selfcheck():
    return B.build_selfcheck(verbose=False)


def _clean(out):
    return "AddressSanitizer" not in out and "runtime error" not in out and "LeakSanitizer" not in out


def test_generated_meshes_and_operators(selfcheck):
    rc, out = B.run_selfcheck()
    assert rc == 0 and "selfcheck: ok" in out and _clean(out), out


@pytest.mark.parametrize("dtype", ["f8", "f4"])
def test_datain_directories(selfcheck, dtype):
    pb = uniform_problem(3, 4, 4, 4, bc={"x": "bc1", "y": "pbc", "z": "bc2"})
    with tempfile.TemporaryDirectory() as d:
        O.write_datain(pb, d)
        if dtype == "f4":
            for n in REALS:
                np.fromfile(os.path.join(d, n + ".bin")).astype(np.float32).tofile(os.path.join(d, n + ".bin"))
        rc, out = B.run_selfcheck(d, 3)
        assert rc == 0 and f"{dtype}," in out and _clean(out), out
        assert "0 Dirichlet faces" not in out  # the x faces are tagged 1


def _corrupt(d, pb, how):
    K, Nfp = pb.K, pb.Nfp
    if how == "truncated":
        with open(os.path.join(d, "vmapP.bin"), "r+b") as f:
            f.truncate(100)
    elif how == "missing":
        os.remove(os.path.join(d, "tz.bin"))
    elif how == "mixed":
        np.fromfile(os.path.join(d, "nx.bin")).astype(np.float32).tofile(os.path.join(d, "nx.bin"))
    elif how == "negative_K":
        np.array([-6], np.int32).tofile(os.path.join(d, "dim.bin"))
    elif how == "huge_K":
        np.array([2**30], np.int32).tofile(os.path.join(d, "dim.bin"))
    elif how == "vmapF_range":
        v = np.fromfile(os.path.join(d, "vmapF.bin"), dtype=np.int32)
        v[5] = 4 * K + 3
        v.tofile(os.path.join(d, "vmapF.bin"))
    elif how == "vmapP_garbage":
        rng = np.random.default_rng(0)
        rng.integers(-2**31, 2**31 - 1, K * 4 * Nfp, dtype=np.int32).tofile(os.path.join(d, "vmapP.bin"))
    elif how == "wrong_order":
        pass  # read for another order below


@pytest.mark.parametrize("how", ["truncated", "missing", "mixed", "negative_K", "huge_K", "vmapF_range",
                                 "vmapP_garbage", "wrong_order"])
def test_corrupted_datain_is_a_clean_error(selfcheck, how):
    pb = cube_problem(3, 2)
    with tempfile.TemporaryDirectory() as d:
        O.write_datain(pb, d)
        _corrupt(d, pb, how)
        rc, out = B.run_selfcheck(d, 4 if how == "wrong_order" else 3)
        assert rc == 2 and "selfcheck:" in out and _clean(out), (how, rc, out)


def test_launcher_build_use_sanitize(selfcheck, tmp_path):
    """build(use_sanitize=True) runs the sanitized self-check over the work
    directory's datain/ and raises on a corrupted one."""
    from exadglm_amd.launcher import DGLM3DSingleGPU

    dg = DGLM3DSingleGPU(None, str(tmp_path / "w"), 3, 5, 1, {"dtype": "f4"})
    dg.init_mesh({"type": "uniform", "param": [3, 3, 3, 0, 1, 0, 1, 0, 1, {"x": "pbc", "y": "pbc", "z": "pbc"}]},
                 verbose=0)
    dg.make_work_dir()
    dg.save_to_binaries(matrix_order="f", verbose=0)
    dg.build(["euler_nopad_main.cu"], use_sanitize=True, verbose=0)
    os.remove(os.path.join(dg.datain_dir, "nz.bin"))
    with pytest.raises(RuntimeError, match="nz.bin"):
        dg.build(["euler_nopad_main.cu"], use_sanitize=True, verbose=0)
