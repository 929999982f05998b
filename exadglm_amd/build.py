"""Build libexadglm_hip.so in-tree for gfx950.

Host-side C++ (operator tables, mesh) is compiled with g++ without FP
contraction so mesh arrays reproduce the reference's bit for bit; the HIP
translation units (kernels + C ABI) are compiled with hipcc for gfx950 and
linked with them into one shared library next to this file.
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

PKG = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(PKG, "csrc")
BUILD = os.path.join(PKG, "_build")
LIB = os.path.join(PKG, "libexadglm_hip.so")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

HOST_SRCS = ["host/refelem.cpp", "host/mesh.cpp", "host/datain.cpp"]
HIP_SRCS = ["solver.hip", "ws_dirichlet.hip"]

HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]
HOST_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-ffp-contract=off", "-Wall", "-Wno-unused-variable"]


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"build step failed ({r.returncode}): {' '.join(cmd)}\n{r.stdout}")
    return r.stdout


SELFCHECK_SRCS = ["host/selfcheck.cpp", "host/refelem.cpp", "host/mesh.cpp", "host/datain.cpp"]
SELFCHECK = os.path.join(BUILD, "selfcheck_sanitized")
SANITIZE_FLAGS = ["-O1", "-g", "-std=c++17", "-ffp-contract=off", "-fno-omit-frame-pointer",
                  "-fsanitize=address,undefined", "-fno-sanitize-recover=all"]


def build_selfcheck(verbose=True):
    """The host-side C++ (operator tables, mesh, datain reader, connectivity
    packing) with its self-check driver, under AddressSanitizer and UBSan
    (g++ only: no device code is sanitized)."""
    os.makedirs(BUILD, exist_ok=True)
    tmp = f"{SELFCHECK}.tmp{os.getpid()}"  # concurrent builders never expose a half-written binary
    _run(["g++", *SANITIZE_FLAGS, *[os.path.join(CSRC, s) for s in SELFCHECK_SRCS], "-o", tmp], verbose)
    os.replace(tmp, SELFCHECK)
    return SELFCHECK


def run_selfcheck(datain_dir=None, N=None):
    """Run the sanitized self-check (optionally over a datain directory of
    order N); returns (exit status, output)."""
    cmd = [SELFCHECK] + ([datain_dir, str(N)] if datain_dir is not None else [])
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:abort_on_error=0:exitcode=3",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1:exitcode=3")
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, env=env)
    return r.returncode, r.stdout


def build(verbose=True):
    os.makedirs(BUILD, exist_ok=True)
    jobs = []
    objs = []
    for s in HOST_SRCS:
        o = os.path.join(BUILD, os.path.basename(s) + ".o")
        jobs.append(["g++", *HOST_FLAGS, "-c", os.path.join(CSRC, s), "-o", o])
        objs.append(o)
    for s in HIP_SRCS:
        o = os.path.join(BUILD, os.path.basename(s) + ".o")
        jobs.append([HIPCC, *HIP_FLAGS, "-c", os.path.join(CSRC, s), "-o", o])
        objs.append(o)
    with ThreadPoolExecutor(max_workers=4) as ex:
        list(ex.map(lambda c: _run(c, verbose), jobs))
    tmp = LIB + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", *objs, "-o", tmp], verbose)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    build(verbose="-q" not in sys.argv)
