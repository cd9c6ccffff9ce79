"""Host-side checks of the rocprofv3 summary script on synthetic CSV files:
region cutting by marker ids (work-group counts), exclusion of autotuner
trial windows, and the copy-kernel calibration of FETCH_SIZE / WRITE_SIZE."""
import csv
import importlib.util
import io
import json
import os
from contextlib import redirect_stdout

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _mod():
    spec = importlib.util.spec_from_file_location("rocprof_summary",
                                                  os.path.join(ROOT, "dmrjtensor.jl_amd", "rocprof_summary.py"))
    m = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(m)
    return m


GEMM = "void dmrj::k_gemm_f64<32, 32>(dmrj::GemmTile const*, int, dmrj::GemmTask const*, dmrj::GemmSeg const*)"
MARK = "dmrj::k_region_marker(int, double*)"

# (name, work-groups, duration ns): before the region, region with one tuning
# window (marker 3 .. 4), after the region
TRACE = [(GEMM, 100, 999), (MARK, 1, 1), (GEMM, 100, 50000), (MARK, 3, 1), (GEMM, 100, 7777), (GEMM, 100, 8888),
         (MARK, 4, 1), (GEMM, 100, 70000), ("rocsolver::foo(int)", 4, 30000), (MARK, 2, 1), (GEMM, 100, 5)]


def _write_trace(path):
    with open(path, "w", newline="") as f:
        w = csv.writer(f)
        w.writerow(["Kind", "Dispatch_Id", "Kernel_Name", "Start_Timestamp", "End_Timestamp", "Workgroup_Size_X",
                    "Grid_Size_X"])
        t = 1000
        for i, (n, wg, d) in enumerate(TRACE, start=1):
            w.writerow(["KERNEL_DISPATCH", i, n, t, t + d, 64, 64 * wg])
            t += d + 10


def _write_counters(path, counter, per_gemm_kib, extra=()):
    with open(path, "w", newline="") as f:
        w = csv.writer(f)
        w.writerow(["Dispatch_Id", "Grid_Size", "Kernel_Name", "Workgroup_Size", "Counter_Name", "Counter_Value"])
        for i, (n, wg, _) in enumerate(TRACE, start=1):
            if n == MARK:
                w.writerow([i, 64 * wg, n, 64, counter, 0])
            elif "k_gemm_f64" in n:
                # counters split over two rows (per-XCD sums) must be added
                w.writerow([i, 64 * wg, n, 64, counter, per_gemm_kib.get(i, 1e6) / 2])
                w.writerow([i, 64 * wg, n, 64, counter, per_gemm_kib.get(i, 1e6) / 2])
        for row in extra:
            w.writerow(row)


def test_trace_region_excludes_tuning(tmp_path):
    m = _mod()
    p = str(tmp_path / "kt.csv")
    _write_trace(p)
    buf = io.StringIO()
    with redirect_stdout(buf):
        m.trace_summary(p, "k_gemm_f64")
    out = buf.getvalue()
    # region GEMMs: dispatches 3 (50000 ns) and 8 (70000 ns); 5, 6 are tuning, 1 and 11 outside
    assert "2 launches, average 0.0600 ms" in out
    assert "rocsolver::foo | 1 | 30.00 | 1 | 30.00 | 0.030 |" in out


def test_traffic_calibrated(tmp_path):
    m = _mod()
    cal_doubles = 1 << 20
    # the copy kernel moves 8 MiB each way; the counter reads half of that on fetch
    cal_f, cal_w = str(tmp_path / "cf.csv"), str(tmp_path / "cw.csv")
    for path, counter, kib in ((cal_f, "FETCH_SIZE", 4096.0), (cal_w, "WRITE_SIZE", 8192.0)):
        with open(path, "w", newline="") as f:
            w = csv.writer(f)
            w.writerow(["Dispatch_Id", "Grid_Size", "Kernel_Name", "Workgroup_Size", "Counter_Name",
                        "Counter_Value"])
            w.writerow([1, 65536, "dmrj::k_copy(dmrj::CopyOp const*, int)", 256, counter, kib])
    pf, pw = str(tmp_path / "pf.csv"), str(tmp_path / "pw.csv")
    _write_counters(pf, "FETCH_SIZE", {3: 100.0, 8: 300.0})
    _write_counters(pw, "WRITE_SIZE", {3: 10.0, 8: 30.0})

    class A:
        pass
    a = A()
    a.fetch, a.write, a.cal_fetch, a.cal_write = pf, pw, cal_f, cal_w
    a.cal_doubles, a.kernel, a.json, a.meta = cal_doubles, "k_gemm_f64", str(tmp_path / "t.json"), '{"L": 8}'
    with redirect_stdout(io.StringIO()):
        m.traffic_summary(a)
    d = json.load(open(a.json))
    # mean over the two region launches, scaled 2x (fetch) and 1x (write)
    assert abs(d["fetch_bytes_per_launch"] - 200.0 * 1024 * 2) < 1e-6
    assert abs(d["write_bytes_per_launch"] - 20.0 * 1024) < 1e-6
    assert d["launches_fetch"] == 2 and d["L"] == 8


def test_mfma_duration_weighted(tmp_path):
    m = _mod()
    p = str(tmp_path / "mf.csv")
    with open(p, "w", newline="") as f:
        w = csv.writer(f)
        w.writerow(["Dispatch_Id", "Grid_Size", "Kernel_Name", "Workgroup_Size", "Counter_Name", "Counter_Value",
                    "Start_Timestamp", "End_Timestamp"])
        t = 0
        util = {3: 40.0, 8: 10.0}
        for i, (n, wg, d) in enumerate(TRACE, start=1):
            if n == MARK:
                w.writerow([i, 64 * wg, n, 64, "MfmaUtil", 0, t, t + d])
            elif "k_gemm_f64" in n:
                w.writerow([i, 64 * wg, n, 64, "MfmaUtil", util.get(i, 99.0), t, t + d])
                w.writerow([i, 64 * wg, n, 64, "MfmaFlopsF64", 1e9, t, t + d])
                w.writerow([i, 64 * wg, n, 64, "MfmaFlopsF16", 3e9 if i == 3 else 0.0, t, t + d])
            t += d + 10
    buf = io.StringIO()
    with redirect_stdout(buf):
        m.mfma_summary(p, "k_gemm_f64", str(tmp_path / "m.json"))
    d = json.load(open(str(tmp_path / "m.json")))
    # region launches 3 (50000 ns, 40 %) and 8 (70000 ns, 10 %)
    assert d["launches"] == 2
    assert abs(d["mfma_util_pct"] - (40.0 * 50000 + 10.0 * 70000) / 120000) < 1e-12
    assert d["mfma_flops_f64_per_launch"] == 1e9
    assert d["mfma_flops_f16_per_launch"] == 1.5e9


def test_counters_duration_weighted(tmp_path):
    m = _mod()
    p = str(tmp_path / "c.csv")
    with open(p, "w", newline="") as f:
        w = csv.writer(f)
        w.writerow(["Dispatch_Id", "Grid_Size", "Kernel_Name", "Workgroup_Size", "Counter_Name", "Counter_Value",
                    "Start_Timestamp", "End_Timestamp"])
        t = 0
        for i, (n, wg, d) in enumerate(TRACE, start=1):
            if n == MARK:
                w.writerow([i, 64 * wg, n, 64, "VALUBusy", 0, t, t + d])
            elif "k_gemm_f64" in n:
                w.writerow([i, 64 * wg, n, 64, "VALUBusy", 20.0 if i == 3 else 60.0, t, t + d])
                w.writerow([i, 64 * wg, n, 64, "MemUnitStalled", 5.0, t, t + d])
            t += d + 10
    buf = io.StringIO()
    with redirect_stdout(buf):
        m.counters_summary(p, "k_gemm_f64", str(tmp_path / "c.json"))
    d = json.load(open(str(tmp_path / "c.json")))
    # region launches 3 (50000 ns) and 8 (70000 ns)
    assert d["launches"] == 2
    assert abs(d["VALUBusy"] - (20.0 * 50000 + 60.0 * 70000) / 120000) < 1e-12
    assert d["MemUnitStalled"] == 5.0
