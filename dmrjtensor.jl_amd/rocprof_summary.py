"""Summaries of rocprofv3 CSV output for the DMRG benchmark.

The benchmark brackets its timed sweeps with two dispatches of the empty
``k_region_marker`` kernel; this script cuts kernel traces and counter
collections to the dispatches between those markers, so that the figures
describe the same launches as the benchmark's own HIP-event timing.

    python rocprof_summary.py trace KERNEL_TRACE.csv [--kernel k_gemm_f64]
    python rocprof_summary.py traffic --fetch F.csv --write W.csv \
        --cal-fetch CF.csv --cal-write CW.csv --cal-doubles N [--kernel k_gemm_f64] [--json OUT]

``trace`` prints per-kernel launch counts and durations (whole run and the
marked region).  ``traffic`` prints the HBM bytes per launch of one kernel in
the marked region from FETCH_SIZE and WRITE_SIZE (KiB per dispatch), each
scaled by what the same counter reads for the copy kernel of the calibration
run, which moves exactly 8*N bytes each way.
"""
import argparse
import csv
import json
import sys

MARKER = "k_region_marker"


def _col(header, *names):
    low = {h.lower(): h for h in header}
    for n in names:
        if n.lower() in low:
            return low[n.lower()]
    raise KeyError("none of %s in CSV header %s" % (names, header))


def _rows(path):
    with open(path, newline="") as f:
        r = csv.DictReader(f)
        return r.fieldnames, list(r)


def _region(ids_names):
    """(first, last) dispatch ids strictly between the first two markers."""
    marks = sorted(i for i, n in ids_names if MARKER in n)
    if len(marks) < 2:
        return None
    return marks[0], marks[1]


def trace_summary(path, kernel):
    hdr, rows = _rows(path)
    cn, cs, ce = _col(hdr, "Kernel_Name"), _col(hdr, "Start_Timestamp"), _col(hdr, "End_Timestamp")
    cd = _col(hdr, "Dispatch_Id", "Correlation_Id")
    recs = [(int(r[cd]), r[cn], int(r[ce]) - int(r[cs])) for r in rows]
    reg = _region([(i, n) for i, n, _ in recs])
    out = {}
    for scope in ("run", "region"):
        agg = {}
        for i, n, d in recs:
            if scope == "region" and (reg is None or not (reg[0] < i < reg[1])):
                continue
            short = n.split("(")[0].replace("void ", "").strip()
            a = agg.setdefault(short, [0, 0])
            a[0] += 1
            a[1] += d
        out[scope] = agg
    print("| kernel | launches (run) | avg us (run) | launches (region) | avg us (region) | total ms (region) |")
    print("|---|---|---|---|---|---|")
    names = sorted(out["region"] or out["run"], key=lambda k: -(out["region"].get(k, [0, 0])[1]))
    for k in names:
        r, g = out["run"].get(k, [0, 0]), out["region"].get(k, [0, 0])
        print("| %s | %d | %.2f | %d | %.2f | %.3f |" % (k, r[0], r[1] / max(r[0], 1) / 1e3, g[0],
                                                         g[1] / max(g[0], 1) / 1e3, g[1] / 1e6))
    sel = [v for k, v in out["region"].items() if kernel in k]
    if sel:
        n = sum(v[0] for v in sel)
        t = sum(v[1] for v in sel)
        print("\n%s in the marked region: %d launches, average %.4f ms" % (kernel, n, t / n / 1e6))


def _per_dispatch(path, counter):
    hdr, rows = _rows(path)
    cn, cd = _col(hdr, "Kernel_Name"), _col(hdr, "Dispatch_Id", "Correlation_Id")
    cc, cv = _col(hdr, "Counter_Name"), _col(hdr, "Counter_Value")
    vals, names = {}, {}
    for r in rows:
        i = int(r[cd])
        names[i] = r[cn]
        if r[cc].startswith(counter):
            vals[i] = vals.get(i, 0.0) + float(r[cv])
    return vals, names


def _kernel_mean(path, counter, kernel, need_region=True):
    vals, names = _per_dispatch(path, counter)
    reg = _region(names.items())
    if need_region and reg is None:
        raise SystemExit("%s: no pair of %s dispatches" % (path, MARKER))
    sel = [v for i, v in vals.items() if kernel in names[i] and (not need_region or reg[0] < i < reg[1])]
    if not sel:
        raise SystemExit("%s: no %s dispatches with %s" % (path, kernel, counter))
    return sum(sel) / len(sel), len(sel)


def traffic_summary(a):
    cal_bytes = 8.0 * a.cal_doubles
    cf, _ = _kernel_mean(a.cal_fetch, "FETCH_SIZE", "k_copy", need_region=False)
    cw, _ = _kernel_mean(a.cal_write, "WRITE_SIZE", "k_copy", need_region=False)
    sf, sw = cal_bytes / (cf * 1024.0), cal_bytes / (cw * 1024.0)
    f, nf = _kernel_mean(a.fetch, "FETCH_SIZE", a.kernel)
    w, nw = _kernel_mean(a.write, "WRITE_SIZE", a.kernel)
    fb, wb = f * 1024.0 * sf, w * 1024.0 * sw
    print("calibration: copy of %.0f bytes each way reads FETCH_SIZE %.1f KiB (scale %.4f), "
          "WRITE_SIZE %.1f KiB (scale %.4f)" % (cal_bytes, cf, sf, cw, sw))
    print("%s in the marked region: FETCH_SIZE %.1f KiB/launch over %d launches -> %.4g bytes; "
          "WRITE_SIZE %.1f KiB/launch over %d launches -> %.4g bytes; total %.4g bytes/launch"
          % (a.kernel, f, nf, fb, w, nw, wb, fb + wb))
    if a.json:
        d = {"kernel": a.kernel, "fetch_bytes_per_launch": fb, "write_bytes_per_launch": wb,
             "fetch_kib_raw": f, "write_kib_raw": w, "calibration_scale_fetch": sf,
             "calibration_scale_write": sw, "launches_fetch": nf, "launches_write": nw}
        if a.meta:
            d.update(json.loads(a.meta))
        with open(a.json, "w") as fo:
            json.dump(d, fo, indent=1)


def main(argv=None):
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="cmd", required=True)
    t = sub.add_parser("trace")
    t.add_argument("csv")
    t.add_argument("--kernel", default="k_gemm_f64")
    f = sub.add_parser("traffic")
    f.add_argument("--fetch", required=True)
    f.add_argument("--write", required=True)
    f.add_argument("--cal-fetch", required=True)
    f.add_argument("--cal-write", required=True)
    f.add_argument("--cal-doubles", type=float, required=True)
    f.add_argument("--kernel", default="k_gemm_f64")
    f.add_argument("--json")
    f.add_argument("--meta", help="extra JSON fields for the output file")
    a = ap.parse_args(argv)
    if a.cmd == "trace":
        trace_summary(a.csv, a.kernel)
    else:
        traffic_summary(a)


if __name__ == "__main__":
    sys.exit(main())
