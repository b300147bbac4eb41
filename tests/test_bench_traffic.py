"""bench.py's counter-traffic writer on synthetic rocprofv3 csv output: the
copy calibration and the per-class averages over dispatches."""
import csv
import json
import os

import bench


def _write(d, rows):
    os.makedirs(d, exist_ok=True)
    with open(os.path.join(d, "run_counter_collection.csv"), "w", newline="") as f:
        w = csv.DictWriter(f, fieldnames=["Dispatch_Id", "Kernel_Name", "Counter_Name",
                                          "Counter_Value"])
        w.writeheader()
        for r in rows:
            w.writerow(r)


def test_traffic_from_counters(tmp_path):
    kib = 256 * 1024
    # the copy reads half its bytes by FETCH_SIZE (two 32 B halves per 64 B
    # request), all of them by WRITE_SIZE
    _write(tmp_path / "copy_fetch", [{"Dispatch_Id": i, "Kernel_Name": "__amd_rocclr_copyBuffer",
                                      "Counter_Name": "FETCH_SIZE", "Counter_Value": kib / 2}
                                     for i in range(5)])
    _write(tmp_path / "copy_write", [{"Dispatch_Id": i, "Kernel_Name": "__amd_rocclr_copyBuffer",
                                      "Counter_Name": "WRITE_SIZE", "Counter_Value": kib}
                                     for i in range(5)])
    # two dispatches of the class's kernel, averaged
    a = b = f"void sparc::{bench.CLASS_KERNELS['xlines'][0]}(int)"
    fe = [{"Dispatch_Id": 1, "Kernel_Name": a, "Counter_Name": "FETCH_SIZE", "Counter_Value": 100},
          {"Dispatch_Id": 2, "Kernel_Name": b, "Counter_Name": "FETCH_SIZE", "Counter_Value": 300}]
    wr = [{"Dispatch_Id": 1, "Kernel_Name": a, "Counter_Name": "WRITE_SIZE", "Counter_Value": 50},
          {"Dispatch_Id": 2, "Kernel_Name": b, "Counter_Name": "WRITE_SIZE", "Counter_Value": 50}]
    _write(tmp_path / "fetch", fe)
    _write(tmp_path / "write", wr)
    out = tmp_path / "traffic.json"
    bench.write_traffic(str(tmp_path), str(out), "box2:fp32", {"xlines": (1.0, 4.0e5, 2)})
    t = json.loads(out.read_text())
    assert abs(t["_calibration"]["fetch_scale"] - 2.0) < 1e-12
    assert abs(t["_calibration"]["write_scale"] - 1.0) < 1e-12
    e = t["box2:fp32:xlines"]
    assert e["dispatches"] == 2
    assert abs(e["fetch_bytes"] - 200 * 1024 * 2.0) < 1e-6
    assert abs(e["write_bytes"] - 50 * 1024) < 1e-6
    assert e["algorithmic_bytes_per_launch"] == 2.0e5
