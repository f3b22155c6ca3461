"""Host-side logic of bench.py (no device): the CPU baseline legs, the committed traffic summaries and
the one-GPU contract."""
import json
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def bench():
    sys.path.insert(0, ROOT)
    import bench as b
    return b


def test_cpu_baseline_single_hierarchy_runs_both_legs_and_names_the_faster(bench):
    insts, dt, _ = bench.instances("c1", 1)
    r = bench.cpu_baseline(insts, dt, r0_budget_s=0.3, r1_budget_s=0.2, max_steps=1000, workers=2)
    r0, r1 = r["R0"], r["R1"]
    assert r0["steps"] >= 2 and r1["steps"] >= 2  # a single hierarchy fills its budget with steps
    assert r0["steps"] <= 1000 and r1["steps"] <= 1000
    assert r0["ado_updates"] == 21 * r0["steps"] and r1["ado_updates"] == 21 * r1["steps"]
    assert r["value"] == max(r0["value"], r1["value"])
    assert r["leg"] in ("R0", "R1") and r["sample"].startswith(r["leg"])
    assert r["cores"] == (1 if r["leg"] == "R0" else r1["cores"])


def test_cpu_baseline_caps_steps_at_the_workload_count(bench):
    insts, dt, _ = bench.instances("c1", 1)
    r = bench.cpu_baseline(insts, dt, r0_budget_s=5.0, r1_budget_s=5.0, max_steps=3, workers=2)
    assert r["R0"]["steps"] == 3 and r["R1"]["steps"] == 3


def test_measured_traffic_attaches_only_for_the_recorded_tiling(bench):
    path = bench.TRAFFIC_FILES["c3"]
    with open(path) as f:
        t = json.load(f)
    tiling = dict(t["tiling"], classes=t["classes"])
    name = "heom_proj_kernel<float, float, float, 3, true, true, true, 1, false>"
    want = t["per_dispatch"][name]["fetch_bytes"] + t["per_dispatch"][name]["write_bytes"]
    assert bench.measured_traffic(name, "c3", t["workload"], t["packed"], tiling) == want
    other = dict(tiling, target_elems=tiling["target_elems"] * 2)
    assert bench.measured_traffic(name, "c3", t["workload"], t["packed"], other) is None
    assert bench.measured_traffic(name, "c3", t["workload"] + " (other)", t["packed"], tiling) is None
    assert bench.measured_traffic(name, "c3", t["workload"], not t["packed"], tiling) is None
    cls = [list(c) if c else c for c in tiling["classes"]]
    cls[2] = [cls[2][0], cls[2][1] + 4]
    assert bench.measured_traffic(name, "c3", t["workload"], t["packed"], dict(tiling, classes=cls)) is None


def test_more_than_one_gpu_exits_with_a_message():
    p = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120, cwd=ROOT)
    assert p.returncode == 2 and "one GPU" in p.stderr and p.stdout == ""


def test_accuracy_sample_covers_every_stratum_and_the_largest_instances(bench):
    insts, _, _ = bench.instances("c3", 4096)
    idx = bench.error_sample(insts, n=64, seed=11)
    assert len(idx) == 64 and len(set(idx)) == 64 and idx == sorted(idx)
    strata = {(insts[i]["H"].shape[0], insts[i]["L"]) for i in idx}
    assert len(strata) == 45 == len(bench._strata(insts))
    dmax = max(p["H"].shape[0] for p in insts)
    lmax = max(p["L"] for p in insts if p["H"].shape[0] == dmax)
    big = [i for i, p in enumerate(insts) if p["H"].shape[0] == dmax and p["L"] == lmax]
    assert set(big[:4]) <= set(idx)
    assert idx == bench.error_sample(insts, n=64, seed=11)  # seeded
