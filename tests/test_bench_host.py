"""CPU checks of bench.py's contract: one JSON line with the fields its
readers parse, one GPU only, and the host record of the CPU baseline."""
import json
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)


def test_more_than_one_gpu_is_refused_before_any_device_use():
    p = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert p.returncode == 2
    assert "one GPU" in p.stderr and p.stdout == ""


def test_committed_bench_line_has_the_contract_fields():
    with open(os.path.join(ROOT, "profiles", "bench_thermal_64x64x32.json")) as fh:
        lines = [ln for ln in fh.read().splitlines() if ln.strip()]
    assert len(lines) == 1
    d = json.loads(lines[0])
    for key in ("metric", "value", "unit", "n_gpus", "steps", "warmup", "ms_per_step",
                "higher_is_better", "scaling", "vs_baseline", "dtype", "data", "config"):
        assert key in d, key
    assert d["n_gpus"] == 1 and d["unit"] == "s" and d["higher_is_better"] is False
    assert abs(d["value"] * 1e3 - d["ms_per_step"]) <= 1e-9 * d["ms_per_step"]
    assert "workload" in d["config"] and "model" not in d["config"]
    r = d["roofline"]
    assert r["bound"] == "hbm" and r["unit"] == "GB/s" and r["peak"] == 8000.0
    assert abs(r["frac"] - r["achieved"] / r["peak"]) < 1e-12 and 0.0 < r["frac"] <= 1.0
    assert r["traffic"] is None or r["traffic"] > 0
    c = d["cpu_baseline"]
    assert c["kind"] in ("port", "reference") and c["cores"] >= 1 and c["value"] > 0
    assert c["sample"] and set(c["host"]) >= {"cpu_model", "nproc", "usable_cpus"}
    # the mixed step is the headline value; both precisions stop on rel-res 1e-8
    assert d["value"] == d["mixed"]["step_s"]
    assert d["mixed"]["rel_res"] < 1e-8 and d["fp64"]["rel_res"] < 1e-8
    m = d["accuracy_matched_rel_l2_1e-10"]
    assert m["fp64"]["rel_l2_vs_closed_form"] <= 1e-10
    assert m["mixed"]["rel_l2_vs_closed_form"] <= 1e-10


def test_host_cpu_record():
    import bench
    h = bench.host_cpu()
    assert isinstance(h["cpu_model"], str) and h["cpu_model"]
    assert h["nproc"] == os.cpu_count()
    assert h["usable_cpus"] is None or 1 <= h["usable_cpus"] <= h["nproc"]
