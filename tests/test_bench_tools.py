"""The counter summary behind bench.py's roofline.traffic: per-class HBM bytes per launch from
two rocprofv3 --pmc CSVs, each counter scaled by what it reads for the closing 1 GiB copy."""
import csv
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import bench  # noqa: E402

FIELDS = ["Correlation_Id", "Dispatch_Id", "Kernel_Name", "Counter_Name", "Counter_Value"]


def write_csv(path, counter, rows):
    with open(path, "w", newline="") as f:
        w = csv.DictWriter(f, fieldnames=FIELDS)
        w.writeheader()
        for i, (name, value) in enumerate(rows):
            w.writerow({"Correlation_Id": i + 1, "Dispatch_Id": i + 1, "Kernel_Name": name,
                        "Counter_Name": counter, "Counter_Value": value})


def test_counter_summary_scales_by_the_copy(tmp_path):
    panel = "void (anonymous namespace)::lu_panel_kernel<float>(PanelArgs<float>)"
    gemv = "(anonymous namespace)::gemv_partial_kernel(long, double const*)"
    finish = "(anonymous namespace)::residual_finish_kernel(long, int)"
    copy = "__amd_rocclr_copyBuffer"
    # FETCH_SIZE reads half of the 1 GiB copy (scale 2), WRITE_SIZE all of it (scale 1)
    fetch = [(panel, 1000.0), (panel, 3000.0), (gemv, 500.0), (finish, 10.0), (copy, 524288.0)]
    write = [(panel, 4000.0), (panel, 0.0), (gemv, 20.0), (finish, 4.0), (copy, 1048576.0)]
    fp, wp = tmp_path / "f.csv", tmp_path / "w.csv"
    write_csv(fp, "FETCH_SIZE", fetch)
    write_csv(wp, "WRITE_SIZE", write)
    text = bench.counters_summary(str(fp), str(wp))
    lines = dict(l.split(" = ") for l in text.splitlines() if l.startswith("traffic_bytes_per_launch"))
    # panel: mean fetch 2000 KiB x 2 + mean write 2000 KiB x 1 = 6000 KiB per launch
    assert int(lines["traffic_bytes_per_launch[lu_panel]"]) == 6000 * 1024
    # residual: the helper kernel's counters add to its class launch (one gemv dispatch)
    assert int(lines["traffic_bytes_per_launch[residual_f64]"]) == ((500 + 10) * 2 + (20 + 4)) * 1024
    assert "traffic_bytes_per_launch[lpgemm]" not in lines  # no dispatches of that class


def test_traffic_from_profile_reads_the_committed_summary():
    # the committed summary carries the dominant kernel's figure; unknown classes give None
    assert bench.traffic_from_profile("lu_panel") > 0
    assert bench.traffic_from_profile("no_such_class") is None
