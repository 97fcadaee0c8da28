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
    assert a.modes.split(",")[0] == "mg" and "learned" in a.modes.split(",")
    assert a.steps >= 1 and a.warmup >= 0


@pytest.mark.parametrize("degree", [4, 5])
def test_kernel_names_and_bytes(degree):
    names = bench.mg_kernel_names(degree)
    assert names["post_level0"].startswith(f"k_mg_post<{degree}, 8, 4>")
    assert names["pre_level0"].startswith(f"k_mg_pre<{degree}, 8, 4>")
    # pass B: z (FP32), Az, p, q, r, x read (FP64), mask, p, q, r, x written, next input (FP32)
    assert bench.MG_KERNEL_BYTES["fcg_b"] == 4 + 8 + 4 * 8 + 1 + 4 * 8 + 4 == 81
    assert bench.MG_KERNEL_BYTES["fcg_a"] == 45


def _profile_traffic(name):
    """kernel -> traffic bytes from the HBM-traffic table of a committed
    profile summary: FETCH_SIZE KiB x 2.0 + WRITE_SIZE KiB x 1.0 (the copy
    calibration the summary states), and the traffic MB column it prints."""
    rows = {}
    text = open(os.path.join(ROOT, "profiles", name)).read()
    section = text.split("## HBM traffic per dispatch", 1)[1].split("\n## ", 1)[0]
    for line in section.splitlines():
        cells = [c.strip() for c in line.strip().strip("|").split("|")]
        if len(cells) < 4 or not cells[0].startswith("`"):
            continue
        try:
            fetch, write = (float(c.replace(",", "")) for c in cells[1:3])
        except ValueError:
            continue
        rows[cells[0]] = ((2.0 * fetch + write) * 1024, float(cells[3].split()[0]) * 1e6)
    return rows


def test_mg_traffic_constants_come_from_the_committed_profile():
    """The traffic the bench line reports for the multigrid kernels is the
    one the committed rocprofv3 summary measured, and that summary's counters
    add up to its own traffic column."""
    rows = _profile_traffic("k_mg_spherical_2048.md")
    names = {"fcg_b": "`k_mg_fcg_b`", "fcg_a": "`k_mg_fcg_a`",
             "post_level0": "`k_mg_post<5, 8, 4>`, 2048²", "pre_level0": "`k_mg_pre<5, 8, 4>`, 2048²"}
    for k, row in names.items():
        counted, printed = rows[row]
        assert abs(counted - printed) <= 0.002 * printed, (k, counted, printed)
        assert bench.MG_TRAFFIC_2048[k] == pytest.approx(printed, rel=1e-6), k


def test_sweep_traffic_constants_come_from_the_committed_profile():
    """The traffic the bench line reports for the Jacobi sweeps (plain FP64,
    mixed FP32) is the committed rocprofv3 summary's for the kernel the bench
    names at 2048^2."""
    rows = _profile_traffic("k_cg_march_spherical_2048.md")
    for mode, fp64 in (("plain", True), ("mixed", False)):
        name = bench.sweep_kernel(2048, fp64)
        assert name.startswith("k_cg_march<")
        counted, _ = rows["`" + name.replace(", 24, 6", "") + "`"]
        assert bench.TRAFFIC_2048[mode] == pytest.approx(counted, rel=2e-3), mode
    assert bench.sweep_kernel(4096, True) == "k_cg_sweep<double>"
