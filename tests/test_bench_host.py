"""CPU checks of bench.py's host logic: the CPU-baseline leg on a tiny time
sample, the rocprofv3 counter summary with its copy calibration, and the
single-GPU guard."""
import csv
import json
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
import bench  # noqa: E402


@pytest.mark.parametrize("wl", ["propagate_100_b4", "scatter_1024"])
def test_cpu_baseline_runs(wl):
    r = bench.cpu_baseline(wl, budget_s=0.2)
    assert r["value"] > 0 and r["kind"] == "port" and r["cores"] >= 1
    assert "steps" in r["sample"]


def _write_counters(d, counter, rows):
    os.makedirs(d, exist_ok=True)
    with open(os.path.join(d, "run_counter_collection.csv"), "w", newline="") as f:
        w = csv.DictWriter(f, fieldnames=["Dispatch_Id", "Kernel_Name", "Counter_Name",
                                          "Counter_Value"])
        w.writeheader()
        for i, (k, v) in enumerate(rows):
            w.writerow({"Dispatch_Id": i, "Kernel_Name": k, "Counter_Name": counter,
                        "Counter_Value": v})


def test_summarize_pmc_calibrates_on_largest_copy(tmp_path, monkeypatch):
    cp = "__amd_rocclr_copyBuffer"
    kn = "void wt::k_adj_cs<1024, 8>(wt::AdjArgs)"
    # a 1 MiB calibration copy reads as 512 KiB (scale 2) next to small copies
    _write_counters(tmp_path / "f", "FETCH_SIZE", [(cp, 4.0), (cp, 512.0), (kn, 100.0),
                                                   (kn, 300.0)])
    _write_counters(tmp_path / "w", "WRITE_SIZE", [(cp, 2.0), (cp, 256.0), (kn, 50.0),
                                                   (kn, 50.0)])
    (tmp_path / "f" / "bench_meta.json").write_text(json.dumps({"tuning": {"a": 1},
                                                                "workload": "unit_test"}))
    monkeypatch.setattr(bench, "traffic_file", lambda w: str(tmp_path / f"traffic_{w}.json"))
    bench.summarize_pmc(str(tmp_path / "f"), str(tmp_path / "w"), 2 ** 20)
    d = json.loads((tmp_path / "traffic_unit_test.json").read_text())
    assert d["calibration"]["fetch_scale"] == pytest.approx(2.0)
    assert d["calibration"]["write_scale"] == pytest.approx(4.0)
    k = d["kernels"]["wt::k_adj_cs<1024, 8>"]
    assert k["fetch_bytes"] == pytest.approx(200.0 * 1024 * 2)
    assert k["write_bytes"] == pytest.approx(50.0 * 1024 * 4)
    assert bench.traffic("unit_test", "wt::k_adj_cs<1024, 8>", {"a": 1}) == \
        pytest.approx(k["fetch_bytes"] + k["write_bytes"])
    assert bench.traffic("unit_test", "wt::k_adj_cs<1024, 8>", {"a": 2}) is None


def test_summarize_pmc_tail_dispatches(tmp_path, monkeypatch):
    """With `tail_dispatches` in the bench's meta file a kernel's summary
    covers only its last n dispatches (in dispatch order), and a kernel with
    n = 0 is left out."""
    cp = "__amd_rocclr_copyBuffer"
    ka = "void wt::k_adj_cs_batch<1024, 10>(wt::BatchAdj)"
    kf = "void wt::k_fwd_cs_batch<1024, 10, 0, false>(wt::BatchFwd)"
    # rows out of dispatch order: ids 0..5; the adjoint's last two are 300, 500
    rows = [(cp, 512.0), (ka, 100.0), (ka, 500.0), (kf, 7.0), (ka, 300.0)]
    order = [0, 1, 5, 2, 3]

    def write(d, counter, vals):
        os.makedirs(d, exist_ok=True)
        with open(os.path.join(d, "run_counter_collection.csv"), "w", newline="") as f:
            w = csv.DictWriter(f, fieldnames=["Dispatch_Id", "Kernel_Name", "Counter_Name",
                                              "Counter_Value"])
            w.writeheader()
            for i, (k, v) in zip(order, vals):
                w.writerow({"Dispatch_Id": i, "Kernel_Name": k, "Counter_Name": counter,
                            "Counter_Value": v})

    write(tmp_path / "f", "FETCH_SIZE", rows)
    write(tmp_path / "w", "WRITE_SIZE", [(k, v / 2) for k, v in rows])
    (tmp_path / "f" / "bench_meta.json").write_text(json.dumps({
        "tuning": {"a": 1}, "workload": "unit_tail",
        "tail_dispatches": {"wt::k_adj_cs_batch<1024, 10>": 2,
                            "wt::k_fwd_cs_batch<1024, 10, 0, false>": 0}}))
    monkeypatch.setattr(bench, "traffic_file", lambda w: str(tmp_path / f"traffic_{w}.json"))
    bench.summarize_pmc(str(tmp_path / "f"), str(tmp_path / "w"), 2 ** 20)
    d = json.loads((tmp_path / "traffic_unit_tail.json").read_text())
    k = d["kernels"]["wt::k_adj_cs_batch<1024, 10>"]
    assert k["dispatches"] == 2
    assert k["fetch_kib"] == pytest.approx((300.0 + 500.0) / 2)
    assert "wt::k_fwd_cs_batch<1024, 10, 0, false>" not in d["kernels"]


def test_roofline_takes_the_binding_unit(tmp_path, monkeypatch):
    """roofline(): HBM fraction from the counted bytes, VALU fraction from
    SQ_INSTS_VALU x 64 lanes against 256 CUs x 64 lanes x 2.4 GHz; `frac`
    and `bound` are the larger one; the algorithmic figure sits apart under
    effective_bw; with no counters for the plan there is no fraction."""
    _write_counters(tmp_path / "f", "FETCH_SIZE", [("copy", 1024.0), ("void wt::k(x)", 400.0)])
    _write_counters(tmp_path / "w", "WRITE_SIZE", [("copy", 1024.0), ("void wt::k(x)", 100.0)])
    _write_counters(tmp_path / "v", "SQ_INSTS_VALU", [("void wt::k(x)", 3.0e6),
                                                      ("void wt::k(x)", 5.0e6)])
    (tmp_path / "f" / "bench_meta.json").write_text(json.dumps({"tuning": {"a": 1},
                                                                "workload": "unit_test"}))
    monkeypatch.setattr(bench, "traffic_file", lambda w: str(tmp_path / f"traffic_{w}.json"))
    bench.summarize_pmc(str(tmp_path / "f"), str(tmp_path / "w"), 2 ** 20, [str(tmp_path / "v")])
    avg_s = 40e-6
    r = bench.roofline("unit_test", "wt::k", {"a": 1}, 2, 1.0e7, avg_s, 10)
    hbm = (400.0 + 100.0) * 1024 / avg_s / bench.PEAK_HBM
    valu = 4.0e6 * 64 / avg_s / (256 * 64 * 2.4e9)
    assert r["hbm"]["frac"] == pytest.approx(hbm, rel=1e-3)
    assert r["valu"]["frac"] == pytest.approx(valu, rel=1e-3)
    assert r["bound"] == ("valu" if valu > hbm else "hbm")
    assert r["frac"] == pytest.approx(max(hbm, valu), rel=1e-3) and r["frac"] <= 1.0
    assert r["effective_bw"]["frac_of_hbm_peak"] == pytest.approx(16 * 1e7 / avg_s / 8e12,
                                                                  rel=1e-3)
    none = bench.roofline("unit_test", "wt::k", {"a": 2}, 2, 1.0e7, avg_s, 10)
    assert none["frac"] is None and none["traffic"] is None


def test_multi_gpu_request_exits_with_message():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "4"],
                       capture_output=True, text=True)
    assert r.returncode == 2 and "one GPU" in r.stderr
