"""Summaries of rocprofv3 CSV output for the DMRG benchmark.

The library dispatches the empty ``k_region_marker`` kernel with as many
work-groups as its id: 1 and 2 bracket the benchmark's timed sweeps, 3 and 4
bracket each set of autotuner trial launches, 5 and 6 the headline
benchmark's profiled FP64 sweep, 7 and 8 its profiled mixed sweep.  This
script cuts kernel traces and counter collections to the dispatches between
two markers (--region, default 1 2), less those between a 3 and the next 4,
so that the figures describe the same launches as the benchmark's own
HIP-event timing.

    python rocprof_summary.py trace KERNEL_TRACE.csv [--kernel k_gemm_f64]
    python rocprof_summary.py traffic --fetch F.csv --write W.csv \
        --cal-fetch CF.csv --cal-write CW.csv --cal-doubles N [--kernel k_gemm_f64] [--json OUT]
    python rocprof_summary.py mfma COUNTERS.csv [--kernel k_gemm_f64] [--json OUT]
    python rocprof_summary.py counters COUNTERS.csv [--kernel k_gemm_f64] [--json OUT]

``trace`` prints per-kernel launch counts and durations (whole run and the
marked region).  ``mfma`` reads a run with ``--pmc MfmaUtil MfmaFlopsF64 [MfmaFlopsF16]``
(rocprofv3's derived metrics: MFMA-busy cycles over GUI-active cycles of all
SIMDs, and executed FP64 / FP16 MFMA flops) and prints the duration-weighted
MFMA utilisation and the executed MFMA flops per launch of one kernel.  ``counters``
prints the duration-weighted mean of every counter or derived metric of one
kernel over the marked region.  ``traffic`` prints the HBM bytes per launch of one kernel in
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


def _marker_id(r, hdr):
    """Marker id = work-groups of the dispatch (grid size / work-group size)."""
    g = int(float(r[_col(hdr, "Grid_Size_X", "Grid_Size")]))
    w = int(float(r[_col(hdr, "Workgroup_Size_X", "Workgroup_Size")]))
    return g // max(1, w)


REGION = (1, 2)   # marker ids that open and close the region (--region)


def _selector(markers):
    """markers: sorted (dispatch id, marker id).  Returns a predicate on
    dispatch ids that is true inside the region (markers REGION) and outside
    tuning windows, or None when the region markers are missing."""
    start = next((i for i, m in markers if m == REGION[0]), None)
    if start is None:
        return None
    end = next((i for i, m in markers if m == REGION[1] and i > start), None)
    if end is None:
        return None
    windows, open_at = [], None
    for i, m in markers:
        if m == 3 and open_at is None:
            open_at = i
        elif m == 4 and open_at is not None:
            windows.append((open_at, i))
            open_at = None
    if open_at is not None:
        windows.append((open_at, float("inf")))

    def inside(i):
        return start < i < end and not any(a < i < b for a, b in windows)
    return inside


def trace_summary(path, kernel):
    hdr, rows = _rows(path)
    cn, cs, ce = _col(hdr, "Kernel_Name"), _col(hdr, "Start_Timestamp"), _col(hdr, "End_Timestamp")
    cd = _col(hdr, "Dispatch_Id", "Correlation_Id")
    recs = [(int(r[cd]), r[cn], int(r[ce]) - int(r[cs])) for r in rows]
    markers = sorted((int(r[cd]), _marker_id(r, hdr)) for r in rows if MARKER in r[cn])
    inside = _selector(markers)
    out = {}
    for scope in ("run", "region"):
        agg = {}
        for i, n, d in recs:
            if scope == "region" and (inside is None or not inside(i)):
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
    total = sum(v[1] for v in out["region"].values())
    sel = [v for k, v in out["region"].items() if kernel in k]
    if sel:
        n = sum(v[0] for v in sel)
        t = sum(v[1] for v in sel)
        print("\n%s in the marked region (autotuner trial launches excluded): %d launches, average %.4f ms, "
              "total %.3f ms of %.3f ms kernel time" % (kernel, n, t / n / 1e6, t / 1e6, total / 1e6))


def _per_dispatch(path, counter):
    hdr, rows = _rows(path)
    cn, cd = _col(hdr, "Kernel_Name"), _col(hdr, "Dispatch_Id", "Correlation_Id")
    cc, cv = _col(hdr, "Counter_Name"), _col(hdr, "Counter_Value")
    vals, names, markers = {}, {}, {}
    for r in rows:
        i = int(r[cd])
        names[i] = r[cn]
        if MARKER in r[cn]:
            markers[i] = _marker_id(r, hdr)
        if r[cc].startswith(counter):
            vals[i] = vals.get(i, 0.0) + float(r[cv])
    return vals, names, sorted(markers.items())


def _kernel_mean(path, counter, kernel, need_region=True):
    vals, names, markers = _per_dispatch(path, counter)
    inside = _selector(markers)
    if need_region and inside is None:
        raise SystemExit("%s: no %s dispatches with ids %d and %d" % (path, MARKER, REGION[0], REGION[1]))
    sel = [v for i, v in vals.items() if kernel in names[i] and (not need_region or inside(i))]
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


def mfma_summary(path, kernel, json_out=None):
    hdr, rows = _rows(path)
    cn, cd = _col(hdr, "Kernel_Name"), _col(hdr, "Dispatch_Id", "Correlation_Id")
    cc, cv = _col(hdr, "Counter_Name"), _col(hdr, "Counter_Value")
    cs, ce = _col(hdr, "Start_Timestamp"), _col(hdr, "End_Timestamp")
    util, dur, names, markers = {}, {}, {}, {}
    flops = {"MfmaFlopsF64": {}, "MfmaFlopsF16": {}}
    for r in rows:
        i = int(r[cd])
        names[i] = r[cn]
        dur[i] = int(r[ce]) - int(r[cs])
        if MARKER in r[cn]:
            markers[i] = _marker_id(r, hdr)
        if r[cc] == "MfmaUtil":
            util[i] = util.get(i, 0.0) + float(r[cv])
        elif r[cc] in flops:
            flops[r[cc]][i] = flops[r[cc]].get(i, 0.0) + float(r[cv])
    inside = _selector(sorted(markers.items()))
    if inside is None:
        raise SystemExit("%s: no %s dispatches with ids %d and %d" % (path, MARKER, REGION[0], REGION[1]))
    sel = [i for i in util if kernel in names[i] and inside(i)]
    if not sel:
        raise SystemExit("%s: no %s dispatches with MfmaUtil" % (path, kernel))
    tdur = sum(dur[i] for i in sel)
    wutil = sum(util[i] * dur[i] for i in sel) / max(1, tdur)
    per = {k: sum(v.get(i, 0.0) for i in sel) for k, v in flops.items()}
    print("%s in the marked region: %d launches, MFMA busy %.1f %% of all SIMD cycles (duration-weighted), "
          "executed FP64 MFMA flops %.4g and FP16 MFMA flops %.4g per launch, %.2f / %.2f TFLOP/s over the "
          "counted kernel time"
          % (kernel, len(sel), wutil, per["MfmaFlopsF64"] / len(sel), per["MfmaFlopsF16"] / len(sel),
             per["MfmaFlopsF64"] / max(1, tdur) / 1e3, per["MfmaFlopsF16"] / max(1, tdur) / 1e3))
    if json_out:
        with open(json_out, "w") as fo:
            json.dump({"kernel": kernel, "launches": len(sel), "mfma_util_pct": wutil,
                       "mfma_flops_f64_per_launch": per["MfmaFlopsF64"] / len(sel),
                       "mfma_flops_f16_per_launch": per["MfmaFlopsF16"] / len(sel),
                       "kernel_ns_total": tdur}, fo, indent=1)


def counters_summary(path, kernel, json_out=None):
    """Duration-weighted mean of every counter (or derived metric) of one kernel
    over its dispatches in the marked region."""
    hdr, rows = _rows(path)
    cn, cd = _col(hdr, "Kernel_Name"), _col(hdr, "Dispatch_Id", "Correlation_Id")
    cc, cv = _col(hdr, "Counter_Name"), _col(hdr, "Counter_Value")
    cs, ce = _col(hdr, "Start_Timestamp"), _col(hdr, "End_Timestamp")
    vals, dur, names, markers = {}, {}, {}, {}
    for r in rows:
        i = int(r[cd])
        names[i] = r[cn]
        dur[i] = int(r[ce]) - int(r[cs])
        if MARKER in r[cn]:
            markers[i] = _marker_id(r, hdr)
        vals.setdefault(r[cc], {})
        vals[r[cc]][i] = vals[r[cc]].get(i, 0.0) + float(r[cv])
    inside = _selector(sorted(markers.items()))
    if inside is None:
        raise SystemExit("%s: no %s dispatches with ids %d and %d" % (path, MARKER, REGION[0], REGION[1]))
    out = {"kernel": kernel}
    for name in sorted(vals):
        sel = [i for i in vals[name] if kernel in names[i] and inside(i)]
        if not sel:
            continue
        tdur = sum(dur[i] for i in sel)
        mean = sum(vals[name][i] * dur[i] for i in sel) / max(1, tdur)
        out[name] = mean
        out["launches"] = len(sel)
        print("%s in the marked region: %s = %.4g (duration-weighted mean over %d launches)"
              % (kernel, name, mean, len(sel)))
    if json_out:
        with open(json_out, "w") as fo:
            json.dump(out, fo, indent=1)


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
    mm = sub.add_parser("mfma")
    mm.add_argument("csv")
    mm.add_argument("--kernel", default="k_gemm_f64")
    mm.add_argument("--json")
    cc = sub.add_parser("counters")
    cc.add_argument("csv")
    cc.add_argument("--kernel", default="k_gemm_f64")
    cc.add_argument("--json")
    for sp in (t, f, mm, cc):
        sp.add_argument("--region", type=int, nargs=2, default=[1, 2], metavar=("OPEN", "CLOSE"),
                        help="k_region_marker ids around the region (1 2: the timed sweeps; the headline "
                             "benchmark's profiled FP64 sweep is 5 6, its profiled mixed sweep 7 8)")
    a = ap.parse_args(argv)
    global REGION
    REGION = tuple(a.region)
    if a.cmd == "trace":
        trace_summary(a.csv, a.kernel)
    elif a.cmd == "mfma":
        mfma_summary(a.csv, a.kernel, a.json)
    elif a.cmd == "counters":
        counters_summary(a.csv, a.kernel, a.json)
    else:
        traffic_summary(a)


if __name__ == "__main__":
    sys.exit(main())
