"""The rocprofv3 summary tool (host logic, synthetic counter tables)."""
import csv
import json
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "dmft-ed_amd", "dmed"))
import rocprof  # noqa: E402

FIELDS = ["Dispatch_Id", "Kernel_Name", "Counter_Name", "Counter_Value"]


def _table(path, counter, rows):
    os.makedirs(path, exist_ok=True)
    with open(os.path.join(path, "x_counter_collection.csv"), "w", newline="") as f:
        w = csv.DictWriter(f, FIELDS)
        w.writeheader()
        for i, (k, v) in enumerate(rows):
            w.writerow({"Dispatch_Id": i, "Kernel_Name": k, "Counter_Name": counter,
                        "Counter_Value": v})


def test_traffic_scales_by_the_copy_kernel(tmp_path):
    dim = 1000
    copy = "void dmed::k_scale_copy<double, double>(double*, double const*, long, double)"
    h64 = "void dmed::k_hxv_col<double, 0, 13>(double const*, ...)"
    h32 = "void dmed::k_hxv_col<float, 0, 13>(float const*, ...)"
    # counter units: FETCH 16 bytes, WRITE 1024 bytes
    _table(str(tmp_path / "f"), "FETCH_SIZE", [(copy, 500.0), (h64, 4000.0), (h64, 4200.0),
                                               (h32, 1000.0)])
    _table(str(tmp_path / "w"), "WRITE_SIZE", [(copy, 8000 / 1024), (h64, 9.0), (h64, 7.0),
                                               (h32, 4.0)])
    out = tmp_path / "t.json"
    rocprof.main(["traffic", "--fetch", str(tmp_path / "f"), "--write", str(tmp_path / "w"),
                  "--dim", str(dim), "--out", str(out), "--md", str(tmp_path / "t.md"),
                  "--cmd", "rocprofv3 ..."])
    r = json.loads(out.read_text())
    assert abs(r["fetch_bytes_per_unit"] - 16.0) < 1e-12
    assert abs(r["write_bytes_per_unit"] - 1024.0) < 1e-9
    assert abs(r["fp64_read_bytes"] - 4100 * 16) < 1e-6
    assert abs(r["fp64_write_bytes"] - 8 * 1024) < 1e-6
    assert abs(r["fp32_bytes_per_launch"] - (1000 * 16 + 4 * 1024)) < 1e-6
    assert r["fp64_algorithmic_bytes"] == 16 * dim
    assert "traffic / algorithmic" in (tmp_path / "t.md").read_text()


def test_l2_hit_rates_per_kernel_form(tmp_path):
    copy = "void dmed::k_scale_copy<double, double>(double*, double const*, long, double)"
    h64 = "void dmed::k_hxv_col<double, 0, 1024, 13, 1>(double const*, ...)"
    l32 = "void dmed::k_hxv_col<float, 1, 1024, 13, 2>(float const*, ...)"
    d = tmp_path / "l2"
    os.makedirs(d)
    with open(d / "x_counter_collection.csv", "w", newline="") as f:
        w = csv.DictWriter(f, FIELDS)
        w.writeheader()
        rows = [(copy, "TCC_HIT_sum", 10.0), (copy, "TCC_MISS_sum", 30.0),
                (h64, "TCC_HIT_sum", 50.0), (h64, "TCC_MISS_sum", 50.0),
                (h64, "TCC_HIT_sum", 70.0), (h64, "TCC_MISS_sum", 30.0),
                (l32, "TCC_HIT_sum", 3.0), (l32, "TCC_MISS_sum", 1.0)]
        for i, (k, c, v) in enumerate(rows):
            w.writerow({"Dispatch_Id": i, "Kernel_Name": k, "Counter_Name": c,
                        "Counter_Value": v})
    out = tmp_path / "l2.md"
    rocprof.main(["l2", "--dir", str(d), "--out", str(out), "--cmd", "rocprofv3 ..."])
    lines = {l.split("|")[2].strip(): l for l in out.read_text().splitlines()
             if l.startswith("| `")}
    assert lines["copy"].rstrip(" |").endswith("0.250")
    assert lines["fp64"].rstrip(" |").endswith("0.600")          # (50+70)/(100+100)
    assert lines["fp32 Lanczos"].rstrip(" |").endswith("0.750")
    assert "fp32" not in lines                                     # no plain FP32 rows
