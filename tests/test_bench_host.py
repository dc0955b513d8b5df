"""bench.py's host-side pieces, without a GPU: the one-GPU refusal, the branch-count sanity check,
the per-unit flop constants against a hand count of one quadratic step, and the committed
rocprofv3 traffic summaries it reads."""
import json
import os
import subprocess
import sys

import numpy as np
import pytest

ROOT = os.path.join(os.path.dirname(__file__), "..")
sys.path.insert(0, ROOT)
import bench  # noqa: E402


def test_more_than_one_gpu_is_refused_with_a_message():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2 and "one GPU" in r.stderr


def test_branch_counts_must_add_up():
    bench.check_counts(np.array([90, 10, 190, 0]), 100)  # 2 normals per quadratic step, 1 per exponential
    with pytest.raises(SystemExit):
        bench.check_counts(np.array([90, 10, 180, 0]), 100)
    with pytest.raises(SystemExit):
        bench.check_counts(np.array([90, 9, 189, 0]), 100)


def test_flops_of_a_quadratic_step():
    # moments 8, update and log-price step 12, Asian sum 1; quadratic branch 14; two central normals
    # of 2 (s) + 2 (t) + 1 (log2) + 18 (9 FMAs) + 2 (1 - 2s) + 1 (product) = 26 each
    assert bench.F_STEP == 8 + 12 + 1
    assert bench.F_QUAD == 14
    assert bench.F_CENTRAL == 2 + 2 + 1 + 18 + 2 + 1 == bench.F_TAIL
    assert bench.F_STEP + bench.F_QUAD + 2 * bench.F_CENTRAL == 87


@pytest.mark.parametrize("workload,name,paths", [("c3", "uneven_batch", 0), ("c2", "grid", 1_000_000)])
def test_traffic_summaries(workload, name, paths):
    d = json.load(open(os.path.join(ROOT, "profiles", f"k_path_fp32_{name}_traffic.json")))
    cal = d["copy_calibration"]
    hbm = 1024 * (d["fetch_size_kib_per_dispatch"] * cal["fetch_scale"] + d["write_size_kib_per_dispatch"] *
                  cal["write_scale"])
    assert abs(hbm - d["hbm_bytes_per_launch"]) <= 1e-6 * hbm
    assert bench.pmc_traffic(workload, paths, {"paths_per_thread": d["paths_per_thread"]}) == d["hbm_bytes_per_launch"]
    # another launch shape or workload size was not profiled
    assert bench.pmc_traffic(workload, paths, {"paths_per_thread": 3}) is None
    assert bench.pmc_traffic(workload, paths + 1, {"paths_per_thread": d["paths_per_thread"]}) is None
