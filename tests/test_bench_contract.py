"""bench.py's command-line contract: one JSON line with the fields the
measurement needs, exit status 0 for a validated run, and a refusal for more
than one GPU or for the serving workload (which is switched off)."""
import json
import subprocess
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
KEYS = {"metric", "value", "unit", "n_gpus", "steps", "warmup", "ms_per_step", "higher_is_better", "scaling",
        "vs_baseline", "dtype", "data", "config", "roofline"}


def run(*args, timeout=600):
    return subprocess.run([sys.executable, str(ROOT / "bench.py"), *args], capture_output=True, text=True,
                          timeout=timeout, cwd=ROOT)


def last_json(stdout: str) -> dict:
    lines = [l for l in stdout.splitlines() if l.startswith("{")]
    assert len(lines) == 1, stdout
    return json.loads(lines[0])


def test_more_than_one_gpu_is_refused():
    r = run("--gpus", "2")
    assert r.returncode == 2 and "one GPU only" in r.stderr


def test_serving_workload_is_refused():
    r = run("--workload", "serve")
    assert r.returncode == 2 and "switched off" in r.stderr


@pytest.mark.gpu
def test_default_workload_line(gpu):
    r = run("--steps", "3", "--warmup", "1", "--no-cpu")
    assert r.returncode == 0, r.stderr[-2000:]
    d = last_json(r.stdout)
    assert KEYS <= set(d)
    assert d["n_gpus"] == 1 and d["steps"] == 3 and d["validated"] is True and d["unit"] == "GB/s"
    roof = d["roofline"]
    assert roof["bound"] == "hbm" and roof["peak"] == 8000.0 and 0 < roof["frac"] < 1
    assert roof["algorithmic_bytes_per_launch"] == 2 * 8 * 30000 ** 2 + 2 * 8 * 29996 ** 2
    assert set(d["hbm_measured"]) == {"copy", "triad", "sweep_mix", "copy_best", "copy_best_shape"}
    rep = d["repeats"]
    assert rep["runs"] == 5 and rep["sweeps_each"] == 3 and len(rep["gbs"]) == 5
    assert sorted(rep["gbs"])[2] == rep["gbs_median"] and rep["gbs_median"] > 0
    # the traffic figure belongs to the tiling that ran, or is null
    assert roof["traffic_tiling"] and (roof["traffic"] is None or
                                       roof["traffic_tiling"] in roof["traffic_profiled_tilings"])
    f32 = d["fp32_storage"]
    assert f32["validated"] is True and f32["points_per_s_vs_f64"] > 1.0
    err = f32["error_vs_f64_seeded"]
    assert 0 < err["normwise_l2"] < 1e-5 and err["target_normwise"] == 1e-6
    assert 0 <= err["l1_norm_rel"] < 1e-6
    bat = d["batched_config4"]
    assert bat["solves_per_s"] > bat["one_by_one_f64_solves_per_s"] > 0
    assert bat["max_rel_norm_diff_vs_f64"] < 1e-5
    # every batched FP32 result has its error against FP64, within the bound
    errs = bat["fp32_vs_fp64_all"]
    assert errs["solves"] == errs["fp32_solves"] == 4096
    assert errs["fp32_within_bound"] == 4096 and 0 < errs["normwise_l2"] < 2e-6


@pytest.mark.gpu
def test_slab_workload_line(gpu):
    r = run("--workload", "star2_f64_slab8_n30000", "--steps", "3", "--no-whole", "--no-tune")
    assert r.returncode == 0, r.stderr[-2000:]
    d = last_json(r.stdout)
    assert KEYS <= set(d) and d["config"]["rows"] == 3750 and d["config"]["halo_rows"] == [2, 2]
