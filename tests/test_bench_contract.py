"""The bench's command-line contract, on the CPU: more than one GPU and
invalid step counts are refused with one JSON line and exit status 2 before
any device work (the roofline's FLOP weights: tests/test_flop_weights.py)."""
import json
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _bench(*args):
    return subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), *args], capture_output=True, text=True,
                          timeout=120, cwd=ROOT)


@pytest.mark.parametrize("n", ["2", "8"])
def test_more_than_one_gpu_is_refused(n):
    r = _bench("--gpus", n)
    assert r.returncode == 2
    line = json.loads(r.stdout.strip().splitlines()[-1])
    assert line["n_gpus"] == int(n) and "one GPU" in line["error"]


@pytest.mark.parametrize("args", [("--steps", "0"), ("--warmup", "-1")])
def test_invalid_step_counts_are_refused(args):
    r = _bench(*args)
    assert r.returncode == 2
    assert "error" in json.loads(r.stdout.strip().splitlines()[-1])


@pytest.mark.gpu
def test_bench_line_carries_the_required_fields(capsys, monkeypatch):
    """A short bench run prints one JSON line with the metric, the roofline of
    the march launch and the CPU baseline, and its headline kernel meets the
    PSNR bar.  Run in this process (no child process is started from one
    that holds the GPU)."""
    sys.path.insert(0, ROOT)
    import bench

    monkeypatch.setattr(sys, "argv", ["bench.py", "--steps", "4", "--warmup", "2", "--light-steps", "200",
                                      "--cpu-seconds", "1", "--no-batch", "--no-4k", "--no-post"])
    assert bench.main() == 0
    line = json.loads(capsys.readouterr().out.strip().splitlines()[-1])
    for key in ("metric", "value", "unit", "n_gpus", "steps", "warmup", "ms_per_step", "higher_is_better", "scaling",
                "vs_baseline", "dtype", "data", "config", "roofline", "cpu_baseline"):
        assert key in line, key
    assert line["n_gpus"] == 1 and line["steps"] == 4 and line["value"] > 0.0
    roof = line["roofline"]
    assert roof["bound"] == "valu" and roof["pipe"] == "valu-fp32" and roof["unit"] == "TFLOP/s"
    assert roof["algorithmic_bytes"] > 0.0 and "traffic_is" in roof
    assert roof["achieved"] > 0.0 and abs(roof["frac"] - roof["achieved"] / roof["peak"]) < 1e-3
    cpu = line["cpu_baseline"]
    assert cpu["value"] > 0.0 and cpu["cores"] >= 1 and cpu["kind"] in ("port", "reference")
    acc = line["accuracy"]
    assert acc["psnr_db_fp64"] >= 45.0
    assert 0.0 <= acc["p999_err_frac_of_peak_fp64"] <= acc["max_err_frac_of_peak_fp64"]
    assert {"x", "y", "rel_vs_fp64", "steps_fp32", "steps_fp64"} <= set(acc["worst_vs_fp32_oracle"])
