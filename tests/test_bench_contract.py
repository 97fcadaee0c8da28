"""bench.py's command-line contract on the CPU: one JSON line refusing
several GPUs (this build measures one), the defaults a bare run uses, and the
per-point byte counts its roofline figures divide by."""
import json
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
import bench  # noqa: E402


def test_several_gpus_are_refused_with_one_json_line():
    p = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=300)
    assert p.returncode == 2
    line = json.loads(p.stdout.strip().splitlines()[-1])
    assert line["n_gpus"] == 2 and "one GPU" in line["error"]


def test_defaults(monkeypatch):
    monkeypatch.setattr(sys, "argv", ["bench.py"])
    a = bench.parse()
    assert (a.gpus, a.n, a.workload, a.relres) == (1, 2048, "solve", 1e-10)
    assert a.modes.split(",")[0] == "learned" and "mg" in a.modes.split(",")
    assert a.steps >= 1 and a.warmup >= 0


@pytest.mark.parametrize("degree", [4, 5])
def test_kernel_names_and_bytes(degree):
    names = bench.mg_kernel_names(degree)
    assert names["post_level0"].startswith(f"k_mg_post<{degree}, 32>")
    assert names["pre_level0"].startswith(f"k_mg_pre<{degree}, 32>")
    # pass B: z (FP32), Az, p, q, r, x read (FP64), mask, p, q, r, x written, next input (FP32)
    assert bench.MG_KERNEL_BYTES["fcg_b"] == 4 + 8 + 4 * 8 + 1 + 4 * 8 + 4 == 81
    assert bench.MG_KERNEL_BYTES["fcg_a"] == 45
    for k, v in bench.MG_TRAFFIC_2048.items():
        # measured traffic within 10% of the algorithmic bytes at 2048^2
        assert 1.0 <= v / (bench.MG_KERNEL_BYTES[k] * 2048 * 2048) <= 1.1
