"""bench.py command-line contract that needs no GPU: more than one GPU is
refused with a message and exit status 2 before any device work."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_bench_refuses_several_gpus():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=300)
    assert r.returncode == 2
    assert "one GPU only" in r.stderr
    assert r.stdout == ""


def test_secondary_leg_failure_is_recorded(monkeypatch, capsys):
    """A failure in one split-BF16 run of the mixed-precision leg is recorded
    in that run's entry (and on stderr); the FP64, FP32 and other split-BF16
    figures are still returned.  run_forward is replaced by a stand-in, so no
    GPU is touched."""
    import importlib.util

    import numpy as np

    spec = importlib.util.spec_from_file_location("bench_under_test", os.path.join(ROOT, "bench.py"))
    bench = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(bench)

    def stand_in(pb, u0, precision="fp64", flux="rusanov", steps=200, warmup=10, kernel_steps=20,
                 stage_kernel="auto"):
        if precision == "bf16x6":
            raise RuntimeError("state is not finite after the timed steps")
        return {"wall": 2.0 if precision == "fp64" else 1.0, "u": np.array(u0), "kernel": precision, "t": 0.0}

    monkeypatch.setattr(bench, "run_forward", stand_in)
    res = bench.mixed_precision_leg(order=1, nhex=1, steps=2, warmup=0)
    assert res["speedup"] == 2.0
    assert res["split_bf16"]["bf16x3"]["kernel"] == "bf16x3"
    assert res["split_bf16"]["bf16x6"] == {"error": "RuntimeError: state is not finite after the timed steps"}
    assert "split-BF16 bf16x6 failed" in capsys.readouterr().err
