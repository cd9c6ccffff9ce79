"""Benchmark: two-site DMRG sweeps on one MI355X.

Default workload (--mode headline, the metric's configuration): Fermi-Hubbard
chain L=64, t=1, U=4, mu=-2 (half filling, particle-hole symmetric), d=4, MPO
bond dimension 6, U(1)xU(1) (N, 2Sz) block-sparse tensors, N=64, Sz=0, seeded
random product-state start (a permutation of 32 up / 32 down singly occupied
sites from numpy default_rng(0)), bond dimension ramped 256 -> 512 -> 1024 ->
2048, two-vector Lanczos (Krylov 2), minm 2, SVD cutoff 0 (the bond dimension
is capped by m alone).

A "step" is one full mixed-precision DMRG sweep (2(L-1) two-site bond steps)
at m = 2048: split-precision H_eff*psi (FP16x3 with per-block scaling by
default, --mixed), FP64 environment updates, FP64 Lanczos scalars and 2x2
Ritz solve on the device, the FP64 Gram-matrix SVD (backend 5), and the FP64
Rayleigh-quotient energy correction at the bond each sweep reports.  The ramp
and --warmup sweeps at m = 2048 are untimed; exactly --steps sweeps are timed
between two stream synchronisations, with all tensors resident in HBM, no
per-kernel events and no autotuner trials inside the timed region.

Then, from one copy each of the resulting state, --cmp-sweeps sweeps of the
plain FP64 drop-in (FP64 everywhere, augmented-matrix SVD, backend 0) and of
the mixed form, both unprofiled: the sweep-time ratio and the relative energy
and centre-bond spectrum deviation of the mixed sweeps from the FP64 ones.
One more sweep of each, with HIP events on every grouped-GEMM launch, gives
the kernels' rooflines.

Output: one JSON line.  value = algorithmic TFLOP/s of the timed mixed
sweeps (H_eff*psi plus environment updates, block-sparse counts,
FP64-equivalent), ms_per_step = mixed sweep wall time.  cpu_baseline = the
reference's whole two-site bond step restated on block-sparse tensors
(oracle/bondstep.py: both H_eff*psi of simplelanczos, the discarded
twosite_update, one LAPACK gesdd per sector, the environment update) on the
host cores at the centre bonds of the same state, extrapolated to a sweep.

Other modes: --mode dmrg (sweeps of one configuration, any model), tdvp,
measure; --compare (precision / SVD variants from one state).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_FP64 = 78.6      # TFLOP/s, MI355X dense FP64 matrix
PEAK_BF16 = 2500.0    # TFLOP/s, MI355X dense BF16 matrix
PEAK_HBM = 8000.0     # GB/s


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--model", default=None, choices=["heisenberg", "hubbard", "j1j2"],
                    help="heisenberg: chain L=100, U(1) Sz (default of --mode dmrg); hubbard: chain L=64, t=1, "
                         "U=4, half filling, U(1)xU(1) (N, Sz) (the headline workload); j1j2: width-4 cylinder, "
                         "J2=0.5, U(1) Sz, D=14 autoMPO")
    ap.add_argument("--L", type=int, default=None, help="sites (default 100 / 64 / 128 by model)")
    ap.add_argument("--m", type=int, default=None, help="bond dimension (default 1024 / 2048 / 1024 by model)")
    ap.add_argument("--ramp", default=None, help="ramp bond dimensions (default 128,256,512 / 256,512,1024)")
    ap.add_argument("--svd-backend", type=int, default=None, help="dmrj_ctx_set_svd_backend mode")
    ap.add_argument("--cutoff", type=float, default=0.0,
                    help="SVD cutoff of the ramp and timed sweeps (0: bond dimension capped by m only)")
    ap.add_argument("--precision", default="f64", choices=["f64", "bf16x3", "f16x3"],
                    help="arithmetic of the H_eff*psi GEMMs")
    ap.add_argument("--env-precision", default=None, choices=["f64", "bf16x3", "f16x3"],
                    help="arithmetic of the environment-update GEMMs (default: same as --precision)")
    ap.add_argument("--max-grow", type=int, default=12)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=25.0,
                    help="budget of the CPU baseline (at least one whole bond step is timed)")
    ap.add_argument("--mixed", default="f16x3", choices=["f16x3", "bf16x3"],
                    help="headline: arithmetic of the mixed form's H_eff*psi GEMMs")
    ap.add_argument("--mixed-svd", type=int, default=5,
                    help="headline: SVD backend of the mixed form (5: FP64 Gram eigensolve)")
    ap.add_argument("--correction", default="sweep", choices=["sweep", "bond"],
                    help="headline: FP64 Rayleigh-quotient energy of the mixed form at the last bond of every "
                         "sweep or at every bond")
    ap.add_argument("--cmp-sweeps", type=int, default=2,
                    help="headline: unprofiled sweeps of the FP64 drop-in and of the mixed form from one state")
    ap.add_argument("--graph", action="store_true",
                    help="capture each Lanczos step (TDVP: each Krylov exponential) into a hipGraph (kernels "
                         "inside graphs are not timed individually, so the roofline then covers the "
                         "environment-update GEMMs only)")
    ap.add_argument("--no-autotune", action="store_true", help="fixed tile-shape heuristic instead of the autotuner")
    ap.add_argument("--tile-order", type=int, default=None, choices=[0, 1],
                    help="grouped-GEMM tile order (dmrj_ctx_set_tile_order; default 1 = tiles sharing a panel "
                         "dealt to one XCD, 0 = plain descending work)")
    ap.add_argument("--eig-batching", type=int, default=None, choices=[0, 1, 2],
                    help="SVD eigensolve batching (dmrj_ctx_set_eig_batching; default 1 = size classes)")
    ap.add_argument("--gemm-cfg", type=int, default=-1,
                    help="force one grouped-GEMM tile shape (dmrj_ctx_set_gemm_cfg, 0..4; -1 = autotuner)")
    ap.add_argument("--traffic-file", default=os.path.join(ROOT, "profiles", "traffic_k_gemm_f64.json"),
                    help="HBM bytes per k_gemm_f64 launch measured with rocprofv3 FETCH_SIZE/WRITE_SIZE")
    ap.add_argument("--method", default="twosite", choices=["twosite", "3S"],
                    help="DMRG method of the timed sweeps (3S: single-site with subspace expansion, FP64)")
    ap.add_argument("--mode", default="headline", choices=["headline", "dmrg", "tdvp", "measure"],
                    help="headline (default): mixed-precision sweeps at m=2048 on the Hubbard chain, then the FP64 "
                         "drop-in and the mixed form from one state; dmrg: two-site DMRG sweeps of one "
                         "configuration; tdvp: real-time two-site TDVP half-sweeps of the Heisenberg chain from "
                         "the Neel state (ComplexF64); measure: after the ramp, time correlationmatrix(psi, S+, "
                         "S-) on the state")
    ap.add_argument("--dt", type=float, default=0.05, help="TDVP time step per half-sweep")
    ap.add_argument("--krylov", type=int, default=12, help="TDVP Krylov dimension")
    ap.add_argument("--compare", action="store_true",
                    help="after the ramp, run --steps sweeps from copies of the same state in each variant "
                         "(FP64 SVD backend 0; FP64 backend 5; BF16x3 and FP16x3 H_eff with FP64 environments, "
                         "backend 5) and report sweep times, speed-ups and |dE/E| against the FP64 drop-in")
    ap.add_argument("--converge", type=int, default=0,
                    help="with --compare: first sweep at m (up to this many sweeps) until the energy changes by "
                         "less than 1e-11 relative, so that the variants start from a converged state")
    ap.add_argument("--progress", action="store_true",
                    help="a progress line every 16 bond steps of two-site sweeps (runs under a counter-collecting "
                         "profiler)")
    ap.add_argument("--step-errors", type=int, default=1,
                    help="headline: 1 = one more mixed sweep whose every bond step is also solved in FP64 and in the "
                         "plain split form from the same inputs (per-bond energy deviations); 0 = skip it")
    ap.add_argument("--calibrate-copy", type=int, default=0,
                    help="only run one block-copy kernel over this many doubles (counter calibration)")
    args = ap.parse_args()
    if args.cmp_sweeps < 1:
        ap.error("--cmp-sweeps must be at least 1")
    if args.gpus != 1:
        print("bench.py measures one GPU: there is no multi-GPU path (one GPU per process; the "
              "sharding of the bond step across GPUs is a design, not code); run with --gpus 1",
              file=sys.stderr)
        sys.exit(2)

    import __graft_entry__ as ge
    pkg = ge.load_package()
    ctx = pkg.default_context()
    if args.calibrate_copy:
        ctx.call("dmrj_selftest_copy", int(args.calibrate_copy))
        log("copy kernel: %d doubles read and written (%.1f MiB each way)"
            % (args.calibrate_copy, args.calibrate_copy * 8 / 2 ** 20))
        return
    if args.no_autotune:
        ctx.set_autotune(0)
    if args.tile_order is not None:
        ctx.set_tile_order(args.tile_order)
    if args.gemm_cfg >= 0:
        ctx.set_gemm_cfg(args.gemm_cfg)
    if args.eig_batching is not None:
        ctx.set_eig_batching(args.eig_batching)
    if args.graph:
        ctx.set_graph(1)
    if args.mode == "tdvp":
        return bench_tdvp(pkg, ctx, args)
    if args.mode == "headline" and not args.compare:
        return bench_headline(pkg, ctx, args)
    if args.model is None:
        args.model = "heisenberg"
    if args.svd_backend is not None:
        ctx.set_svd_backend(args.svd_backend)
    psi, mpo, desc, L, m = make_workload(pkg, ctx, args)
    params = pkg.algvars(psi)

    def sweeps(mm, n, method="twosite"):
        run_sweeps(pkg, psi, mpo, params, mm, n, args.cutoff, precision=args.precision,
                   env_precision=args.env_precision, method=method, progress=args.progress)

    t_ramp = time.perf_counter()
    for mr in [int(x) for x in args.ramp.split(",") if x]:
        sweeps(mr, 1)
        log("ramp m=%d E=%.12f maxm=%d t=%.2fs" % (mr, params.energy, params.history[-1]["biggestm"],
                                                   params.history[-1]["time"]))
    # keep sweeping at the target m until the bond dimension has grown to it
    for _ in range(args.max_grow):
        if params.history and params.history[-1]["biggestm"] >= int(0.95 * m):
            break
        sweeps(m, 1)
        log("grow m=%d E=%.12f maxm=%d t=%.2fs" % (m, params.energy, params.history[-1]["biggestm"],
                                                   params.history[-1]["time"]))
    # the ramp grows the bonds with two-site sweeps; 3S is timed from there
    # with its noise back at the initial 1.0, so that the subspace expansion
    # is active (the noise shrinks by 1% per step while the energy falls and
    # the expansion stops below noise_incr = 0.01, i.e. after ~460 steps)
    if args.method == "3S":
        params.noise = 1.0
    for _ in range(args.warmup):
        sweeps(m, 1, args.method)
        log("warmup m=%d E=%.12f maxm=%d t=%.2fs" % (m, params.energy, params.history[-1]["biggestm"],
                                                     params.history[-1]["time"]))
    log("ramp+warmup %.1fs" % (time.perf_counter() - t_ramp))
    if args.mode == "measure":
        return bench_measure(pkg, ctx, args, psi, desc, L, m)
    if args.compare:
        for k in range(args.converge):
            e_prev = params.energy
            sweeps(m, 1)
            rel = abs(params.energy - e_prev) / abs(params.energy)
            log("converge m=%d E=%.12f dE/E=%.2e t=%.2fs" % (m, params.energy, rel, params.history[-1]["time"]))
            if rel < 1e-11:
                break
        return bench_compare(pkg, ctx, args, psi, mpo, desc, L, m, params)

    ctx.profile(True)
    ctx.sync()
    ctx.flops(reset=True)
    for k in ("gemm_f64", "mix", "copy", "svd", "gemm_lp", "graph", "gemm_f16"):
        ctx.kernel_stats(k, reset=True)
    tune0 = ctx.autotune_stats()
    ctx.svd_stats(reset=True)
    ctx.marker(1)       # kernel-trace cut: start of the timed region
    t0 = time.perf_counter()
    sweeps(m, args.steps, args.method)
    ctx.sync()
    t1 = time.perf_counter()
    ctx.marker(2)       # end of the timed region
    ctx.sync()
    tune1 = ctx.autotune_stats()
    svdst = ctx.svd_stats()
    elapsed = t1 - t0
    fl_heff, fl_env = ctx.flops()
    stats = {k: ctx.kernel_stats(k) for k in ("gemm_f64", "mix", "copy", "svd", "gemm_lp", "graph", "gemm_f16")}
    ctx.profile(False)
    for h in params.history[-args.steps:]:
        log("timed sweep E=%.12f maxtrunc=%.3e m=%d t=%.3fs" % (h["energy"], h["maxtrunc"], h["biggestm"],
                                                               h["time"]))
    sweep_s = elapsed / args.steps
    tflops = (fl_heff + fl_env) / elapsed / 1e12
    # dominant kernel: the grouped GEMM of the arithmetic the sweep ran in
    gkind, gname, gpeak = {"f64": ("gemm_f64", "k_gemm_f64 (FP64 grouped GEMM, MFMA 16x16x4)", PEAK_FP64),
                           "bf16x3": ("gemm_lp", "k_gemm_bf16x3 (BF16x3 split grouped GEMM, MFMA 16x16x32)",
                                      PEAK_BF16),
                           "f16x3": ("gemm_f16", "k_gemm_f16p (scaled FP16x3 grouped GEMM on pre-split "
                                     "operand planes, MFMA 16x16x32)", PEAK_BF16)}[args.precision]
    g = stats[gkind]
    g_tflops = g["flops"] / (g["ms"] * 1e-3) / 1e12 if g["ms"] > 0 else 0.0
    g_avg_ms = g["ms"] / max(1, g["launches"])
    log("%s launches=%d avg=%.4f ms total=%.3f s, %.2f TFLOP/s; mix %.3f s %.0f GB/s; copy %.3f s; "
        "svd %.3f s" % (gkind, g["launches"], g_avg_ms, g["ms"] / 1e3, g_tflops, stats["mix"]["ms"] / 1e3,
                        stats["mix"]["bytes"] / max(1e-9, stats["mix"]["ms"] * 1e-3) / 1e9,
                        stats["copy"]["ms"] / 1e3, stats["svd"]["ms"] / 1e3))

    traffic, traffic_note = None, None
    if os.path.exists(args.traffic_file) and args.precision == "f64" and args.method == "twosite":
        with open(args.traffic_file) as f:
            tr = json.load(f)
        if tr.get("L") == L and tr.get("m") == m and tr.get("kernel") == "k_gemm_f64":
            traffic = tr["fetch_bytes_per_launch"] + tr["write_bytes_per_launch"]
            traffic_note = ("rocprofv3 FETCH_SIZE + WRITE_SIZE per k_gemm_f64 launch in the timed region, "
                            "scaled by the copy-kernel calibration (%s)" % os.path.basename(args.traffic_file))
    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(pkg, psi, mpo, params, args.cpu_seconds, args.method)
    mname = {"twosite": "two-site", "3S": "3S single-site with subspace expansion"}[args.method]
    form = "FP64 GPU drop-in" if args.precision == "f64" else "%s H_eff*psi" % args.precision
    sym = "U(1)xU(1)" if args.model == "hubbard" else "U(1)"
    mpo_name = {"heisenberg": "Heisenberg", "hubbard": "Fermi-Hubbard", "j1j2": "J1-J2 cylinder"}[args.model]
    out = {
        "metric": "DMRG sweep time + H_eff*psi TFLOP/s (%s, %s, %s)" % (form, mname, sym),
        "value": round(tflops, 4),
        "unit": "TFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(sweep_s * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {"f64": "f64", "bf16x3": "bf16x3 (FP32 accumulate, FP64 Lanczos scalars/SVD/energy)",
                  "f16x3": "f16x3 per-block scaled (FP32 accumulate, FP64 Lanczos scalars/SVD/energy)"}[
                      args.precision],
        "data": "synthetic (seeded random product state, %s MPO)" % mpo_name,
        "config": {"workload": "%s, %s DMRG sweep at m=%d (ramp %s), Krylov 2, cutoff %g"
                               % (desc, mname, m, args.ramp, args.cutoff), "model": args.model,
                   "method": args.method, "ramp_method": "twosite",
                   "noise_at_end": params.noise if args.method == "3S" else None,
                   "L": L, "m": m, "energy": params.energy, "precision": args.precision,
                   "env_precision": args.env_precision or args.precision,
                   "svd_backend": args.svd_backend if args.svd_backend is not None else "default",
                   "svd_s_per_step": round(stats["svd"]["ms"] / 1e3 / args.steps, 4),
                   "jacobi_sweeps_per_svd": round(svdst["sweeps"] / svdst["calls"], 2) if svdst["calls"] else None,
                   "heff_flops_per_sweep": fl_heff / args.steps, "env_flops_per_sweep": fl_env / args.steps,
                   "sweep_s": sweep_s},
        "roofline": {"bound": "mfma", "kernel": gname,
                     "achieved": round(g_tflops, 3), "peak": gpeak, "unit": "TFLOP/s",
                     "frac": round(g_tflops / gpeak, 4), "traffic": traffic,
                     "flops_basis": "algorithmic FP64-equivalent flops (split forms execute 3 MFMA products each)",
                     "traffic_note": traffic_note,
                     "avg_launch_ms": round(g_avg_ms, 5), "launches": g["launches"],
                     "flops_per_launch": g["flops"] / max(1, g["launches"]),
                     "compulsory_bytes_per_launch": g["bytes"] / max(1, g["launches"])},
        "autotune": {"keys": tune1["keys"], "tuned_in_timed_region": tune1["tuned"] - tune0["tuned"],
                     "hits_in_timed_region": tune1["hits"] - tune0["hits"]},
        "graph": ctx.graph_stats() if args.graph else None,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)

WORKLOADS = {"heisenberg": (100, 1024, "128,256,512"), "hubbard": (64, 2048, "256,512,1024"),
             "j1j2": (128, 1024, "128,256,512")}


def j1j2_autompo(pkg, Lx, Ly, J1=1.0, J2=0.5):
    """MPO of the J1-J2 cylinder (width Ly, periodic, snaked column-major) by
    autoMPO from its bond list: link dimension 14 at width 4."""
    _, bonds = pkg.j1j2_bulk(Lx, Ly, J1=J1, J2=J2)
    Sp, Sm, Sz, O, Id = pkg.spinOps()
    terms = []
    for (a, b), J in sorted(bonds.items()):
        i, j = min(a, b) + 1, max(a, b) + 1
        terms += [pkg.mpoterm(0.5 * J, Sp, i, Sm, j), pkg.mpoterm(0.5 * J, Sm, i, Sp, j), pkg.mpoterm(J, Sz, i, Sz, j)]
    return pkg.autoMPO(terms, Ns=Lx * Ly, qphys=np.array([[1], [-1]]))


def make_workload(pkg, ctx, args):
    """Seeded product-state start and MPO of one model: (psi, mpo, description,
    L, m).  Fills args.ramp with the model's default ramp when unset."""
    defaults = WORKLOADS[args.model]
    L = args.L or defaults[0]
    m = args.m or defaults[1]
    if args.ramp is None:
        args.ramp = defaults[2]
    rng = np.random.default_rng(0)
    if args.model == "hubbard":
        # half filling, Sz = 0: a seeded permutation of L/2 up and L/2 down singly occupied sites
        states = list(rng.permutation([1] * (L // 2) + [2] * (L - L // 2)))
        qphys = np.array([[0, 0], [1, 1], [1, -1], [2, 0]])
        mpo = pkg.MPO(pkg.makeMPO(pkg.hubbardMPO(), 4, L).H, qphys=qphys)
        psi = pkg.product_state(states, 4, qphys=qphys, ctx=ctx)
        desc = "Fermi-Hubbard L=%d t=1 U=4 mu=-2 d=4 D=6 U(1)xU(1) N=%d Sz=0" % (L, L)
    elif args.model == "j1j2":
        Ly = 4
        Lx = L // Ly
        states = list(rng.permutation([0] * (L // 2) + [1] * (L - L // 2)))
        qphys = np.array([[1], [-1]])
        Ws = j1j2_autompo(pkg, Lx, Ly, J1=1.0, J2=0.5)
        mpo = pkg.MPO(Ws, qphys=qphys)
        psi = pkg.product_state(states, 2, qphys=qphys, ctx=ctx)
        D = max(W.shape[3] for W in Ws)
        desc = "J1-J2 Heisenberg cylinder %dx%d J2=0.5 d=2 D=%d (autoMPO) U(1) Sz=0" % (Lx, Ly, D)
    else:
        states = list(rng.permutation([0] * (L // 2) + [1] * (L - L // 2)))
        qphys = np.array([[1], [-1]])
        mpo = pkg.MPO(pkg.makeMPO(pkg.heisenbergMPO(), 2, L).H, qphys=qphys)
        psi = pkg.product_state(states, 2, qphys=qphys, ctx=ctx)
        desc = "Heisenberg L=%d d=2 D=5 U(1) Sz=0" % L
    return psi, mpo, desc, L, m


def run_sweeps(pkg, psi, mpo, params, m, n, cutoff, precision="f64", env_precision=None, correction=None,
               method="twosite", progress=False):
    """pkg.dmrg(...) with the same settings; with progress, the same sweeps
    through optmps with a line every 16 bond steps (for runs under a
    counter-collecting profiler, where one sweep takes minutes)."""
    if not progress or method != "twosite":
        pkg.dmrg(psi, mpo, m=m, sweeps=n, cutoff=cutoff, params=params, precision=precision,
                 env_precision=env_precision, correction=correction, method=method)
        return
    count = [0]

    def step(nn, j, i, iL, iR, psi_, mpo_, Lenv, Renv, p):
        pkg.twostep(nn, j, i, iL, iR, psi_, mpo_, Lenv, Renv, p)
        count[0] += 1
        if count[0] % 16 == 0:
            log("  bond steps done: %d" % count[0])

    params.method = "twosite"
    params.maxm, params.minm, params.sweeps, params.cutoff = m, 2, n, cutoff
    params.silent, params.goal, params.precision = True, 0.0, precision
    params.env_precision = env_precision
    params.correction = correction or ("none" if precision == "f64" else "sweep")
    pkg.optmps(psi, mpo, params, stepfct=step)


def _free_envs(params):
    for env in (params.Lenv, params.Renv):
        for e in (env.V if env is not None else []):
            if e is not None:
                e.free()
    params.Lenv = params.Renv = None


def _traffic(kernel, model, L, m, precision, sweep):
    """HBM bytes per launch of one grouped-GEMM kernel on this workload (the
    FP64 drop-in sweep or the mixed sweep), from the rocprofv3 FETCH_SIZE /
    WRITE_SIZE summary committed under profiles/ (None when no counter run of
    this workload, sweep and kernel is there)."""
    path = os.path.join(ROOT, "profiles", "traffic_%s_%s_L%d_m%d_%s.json" % (kernel, model, L, m, sweep))
    if not os.path.exists(path):
        return None, None
    with open(path) as f:
        tr = json.load(f)
    if tr.get("kernel") != kernel or tr.get("precision", precision) != precision or tr.get("sweep") != sweep:
        return None, None
    return (tr["fetch_bytes_per_launch"] + tr["write_bytes_per_launch"],
            "rocprofv3 FETCH_SIZE + WRITE_SIZE per %s launch of this workload's profiled sweep, scaled by the "
            "copy-kernel calibration (%s)" % (kernel, os.path.basename(path)))


GEMM_KINDS = {"f64": ("gemm_f64", "k_gemm_f64", "FP64 grouped GEMM, MFMA 16x16x4 f64", PEAK_FP64),
              "bf16x3": ("gemm_lp", "k_gemm_bf16x3", "BF16x3 split grouped GEMM, MFMA 16x16x32 bf16", PEAK_BF16),
              "f16x3": ("gemm_f16", "k_gemm_f16p", "per-block scaled FP16x3 grouped GEMM on pre-split operand planes, "
                        "MFMA 16x16x32 f16; its pre-passes k_region_absmax and k_f16_presplit are timed apart "
                        "(profiled sweep f16_prep_s)", PEAK_BF16)}


def _bound(flops, nbytes, ms, peak_tflops):
    """Roofline placement of a kernel from its algorithmic flops and bytes
    per launch: below the ridge (peak flops / HBM bandwidth) it is HBM-bound
    and measured in GB/s against the HBM peak, above it against the matrix
    peak in TFLOP/s.  Returns (bound, achieved, peak, unit, frac)."""
    ridge = peak_tflops * 1e12 / (PEAK_HBM * 1e9)
    if nbytes > 0 and flops / nbytes < ridge:
        ach = nbytes / (ms * 1e-3) / 1e9
        return "hbm", round(ach, 1), PEAK_HBM, "GB/s", round(ach / PEAK_HBM, 5)
    ach = flops / (ms * 1e-3) / 1e12
    return "mfma", round(ach, 3), peak_tflops, "TFLOP/s", round(ach / peak_tflops, 5)


def _roofline(stats, prec, model, L, m, sweep):
    kind, kname, desc, peak = GEMM_KINDS[prec]
    g = stats[kind]
    if g["launches"] == 0 or g["ms"] <= 0:
        return None
    n = g["launches"]
    bound, ach, pk, unit, frac = _bound(g["flops"] / n, g["bytes"] / n, g["ms"] / n, peak)
    traffic, note = _traffic(kname, model, L, m, prec, sweep)
    return {"bound": bound, "kernel": "%s (%s)" % (kname, desc), "achieved": ach, "peak": pk,
            "unit": unit, "frac": frac, "traffic": traffic,
            "achieved_tflops": round(g["flops"] / (g["ms"] * 1e-3) / 1e12, 3), "matrix_peak_tflops": peak,
            "basis": "algorithmic FP64-equivalent flops and bytes per launch (the FP64 operands and result of "
                     "the contraction; split forms execute 3 MFMA products per FP64-equivalent product, against "
                     "the dense FP16/BF16 matrix peak); bound = hbm when flops/bytes is below the ridge "
                     "peak / 8 TB/s",
            "traffic_note": note, "avg_launch_ms": round(g["ms"] / n, 5), "launches": n,
            "flops_per_launch": g["flops"] / n,
            "compulsory_bytes_per_launch": g["bytes"] / n}


def bench_headline(pkg, ctx, args):
    """Mixed-precision two-site sweeps at m = 2048 on the Hubbard chain (the
    timed steps), then the plain FP64 drop-in and the mixed form from one
    state: sweep-time ratio, |dE/E| and spectrum deviation per sweep."""
    args.model = args.model or "hubbard"
    psi, mpo, desc, L, m = make_workload(pkg, ctx, args)
    mixed = dict(precision=args.mixed, env_precision="f64", correction=args.correction)
    fp64 = dict(precision="f64", env_precision=None, correction=None)
    params = pkg.algvars(psi)
    ctx.set_svd_backend(args.mixed_svd)

    def sweeps(st, p, mm, n, cfg):
        run_sweeps(pkg, st, mpo, p, mm, n, args.cutoff, progress=args.progress, **cfg)

    t_ramp = time.perf_counter()
    for mr in [int(x) for x in args.ramp.split(",") if x]:
        sweeps(psi, params, mr, 1, mixed)
        h = params.history[-1]
        log("ramp m=%d E=%.12f maxm=%d t=%.2fs" % (mr, h["energy"], h["biggestm"], h["time"]))
    for _ in range(args.max_grow):
        if params.history and params.history[-1]["biggestm"] >= int(0.95 * m):
            break
        sweeps(psi, params, m, 1, mixed)
        h = params.history[-1]
        log("grow m=%d E=%.12f maxm=%d t=%.2fs" % (m, h["energy"], h["biggestm"], h["time"]))
    for _ in range(args.warmup):
        sweeps(psi, params, m, 1, mixed)
        h = params.history[-1]
        log("warmup m=%d E=%.12f maxm=%d t=%.2fs" % (m, h["energy"], h["biggestm"], h["time"]))
    log("ramp+warmup %.1fs" % (time.perf_counter() - t_ramp))

    # ---- timed region: --steps mixed sweeps, no events, no autotuner trials
    ctx.set_autotune(3)
    ctx.sync()
    ctx.flops(reset=True)
    tune0 = ctx.autotune_stats()
    ctx.marker(1)
    t0 = time.perf_counter()
    sweeps(psi, params, m, args.steps, mixed)
    ctx.sync()
    t1 = time.perf_counter()
    ctx.marker(2)
    ctx.sync()
    tune1 = ctx.autotune_stats()
    elapsed = t1 - t0
    fl_heff, fl_env = ctx.flops()
    for h in params.history[-args.steps:]:
        log("timed sweep E=%.12f maxtrunc=%.3e m=%d t=%.3fs" % (h["energy"], h["maxtrunc"], h["biggestm"],
                                                               h["time"]))
    _free_envs(params)

    # ---- FP64 drop-in and mixed form from one copy each of the state
    rows = {}
    marks = {"fp64": (5, 6), "mixed": (7, 8)}
    kept = None
    for name, cfg, backend in (("fp64", fp64, 0), ("mixed", mixed, args.mixed_svd)):
        st = psi.copy()
        ctx.set_svd_backend(backend)
        p2 = pkg.algvars(st)
        sweeps(st, p2, m, args.cmp_sweeps, cfg)
        hist = list(p2.history[-args.cmp_sweeps:])
        # one more sweep with HIP events on every grouped-GEMM launch
        ctx.profile(True)
        for k in ("gemm_f64", "mix", "copy", "svd", "gemm_lp", "gemm_f16", "f16_prep"):
            ctx.kernel_stats(k, reset=True)
        ctx.marker(marks[name][0])
        sweeps(st, p2, m, 1, cfg)
        ctx.sync()
        ctx.marker(marks[name][1])
        ctx.sync()
        stats = {k: ctx.kernel_stats(k) for k in ("gemm_f64", "mix", "copy", "svd", "gemm_lp", "gemm_f16", "f16_prep")}
        ctx.profile(False)
        prof = p2.history[-1]
        rows[name] = {"svd_backend": backend, "precision": cfg["precision"],
                      "env_precision": cfg["env_precision"] or cfg["precision"],
                      "sweep_s": [round(h["time"], 4) for h in hist], "energies": [h["energy"] for h in hist],
                      "storeD": hist[-1]["storeD"], "maxtrunc": hist[-1]["maxtrunc"],
                      "biggestm": hist[-1]["biggestm"],
                      "profiled_sweep": {"sweep_s": round(prof["time"], 4), "energy": prof["energy"],
                                         "svd_s": round(stats["svd"]["ms"] / 1e3, 4),
                                         "gemm_f64_s": round(stats["gemm_f64"]["ms"] / 1e3, 4),
                                         "gemm_split_s": round((stats["gemm_lp"]["ms"] + stats["gemm_f16"]["ms"])
                                                               / 1e3, 4),
                                         "f16_prep_s": round(stats["f16_prep"]["ms"] / 1e3, 4),
                                         "mix_s": round(stats["mix"]["ms"] / 1e3, 4)},
                      "_stats": stats}
        log("%s: sweeps %s s, E %s; profiled sweep %.3f s (svd %.3f s, gemm f64 %.3f s, split gemm %.3f s)"
            % (name, rows[name]["sweep_s"], ["%.12f" % e for e in rows[name]["energies"]], prof["time"],
               rows[name]["profiled_sweep"]["svd_s"], rows[name]["profiled_sweep"]["gemm_f64_s"],
               rows[name]["profiled_sweep"]["gemm_split_s"]))
        if name == "fp64":
            kept = (st, p2)
        else:
            _free_envs(p2)
            st.free()
    ctx.set_svd_backend(0)
    ref, mx = rows["fp64"], rows["mixed"]
    t_fp64 = sum(ref["sweep_s"]) / len(ref["sweep_s"])
    t_mx = sum(mx["sweep_s"]) / len(mx["sweep_s"])
    rel = [abs(a - b) / abs(b) for a, b in zip(mx["energies"], ref["energies"])]
    n8 = min(len(mx["storeD"]), len(ref["storeD"]))
    dspec = max([abs(a - b) for a, b in zip(mx["storeD"][:n8], ref["storeD"][:n8])] or [0.0])
    log("FP64 drop-in %.3f s/sweep, mixed %.3f s/sweep: %.2fx; |dE/E| per sweep %s; centre spectrum |dD| %.2e"
        % (t_fp64, t_mx, t_fp64 / t_mx, ["%.2e" % r for r in rel], dspec))

    steperr = None
    if args.step_errors:
        ctx.set_svd_backend(args.mixed_svd)
        st2 = kept[0].copy()
        steperr = bond_step_errors(pkg, ctx, st2, mpo, m, args.cutoff, mixed)
        st2.free()
        ctx.set_svd_backend(0)

    cpu = None
    if not args.no_cpu:
        st, p2 = kept
        ctx.set_svd_backend(0)
        cpu = cpu_sweep_baseline(pkg, ctx, st, mpo, p2, args.cpu_seconds, m, args.cutoff)
    _free_envs(kept[1])
    kept[0].free()

    roof_mx = _roofline(mx["_stats"], args.mixed, args.model, L, m, "mixed")
    roof_64 = _roofline(ref["_stats"], "f64", args.model, L, m, "fp64")
    roof_env = _roofline(mx["_stats"], "f64", args.model, L, m, "mixed")
    for r in rows.values():
        del r["_stats"]
    out = {
        "metric": "DMRG sweep time + H_eff*psi TFLOP/s, mixed (%s H_eff*psi, FP64 env/Lanczos/SVD/energy) vs FP64 "
                  "GPU drop-in, two-site, U(1)xU(1)" % args.mixed,
        "value": round((fl_heff + fl_env) / elapsed / 1e12, 4),
        "unit": "TFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(elapsed / args.steps * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "%s H_eff*psi (FP32 accumulate) + f64 (environments, Lanczos scalars, SVD, energy)" % args.mixed,
        "data": "synthetic (seeded random product state at half filling, Fermi-Hubbard MPO)",
        "config": {"workload": "%s, two-site DMRG sweep at m=%d (ramp %s), Krylov 2, cutoff %g"
                               % (desc, m, args.ramp, args.cutoff),
                   "model": args.model, "L": L, "m": m, "method": "twosite",
                   "mixed": {"precision": args.mixed, "env_precision": "f64", "svd_backend": args.mixed_svd,
                             "energy_correction": args.correction},
                   "energy": params.energy,
                   "heff_flops_per_sweep": fl_heff / args.steps, "env_flops_per_sweep": fl_env / args.steps},
        "speedup_vs_fp64_dropin": round(t_fp64 / t_mx, 3),
        "fp64_dropin_ms_per_sweep": round(t_fp64 * 1e3, 3),
        "mixed_ms_per_sweep_from_same_state": round(t_mx * 1e3, 3),
        "rel_dE_per_sweep": rel,
        "max_rel_dE": max(rel) if rel else None,
        "centre_spectrum_max_abs_dD": dspec,
        "bond_step_rel_dE_vs_fp64": steperr,
        "comparison": rows,
        "roofline": roof_mx,
        "roofline_fp64_dropin": roof_64,
        "roofline_mixed_env_f64": roof_env,
        "autotune": {"keys": tune1["keys"], "tuned_in_timed_region": tune1["tuned"] - tune0["tuned"],
                     "hits_in_timed_region": tune1["hits"] - tune0["hits"]},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


def bond_step_errors(pkg, ctx, st, mpo, m, cutoff, mixed):
    """One mixed sweep from the state st in which every bond step first runs
    the Lanczos step (dmrj_lanczos2) on the same two-site tensor and
    environments in FP64, in the split form, and in the split form with the
    FP64 Rayleigh-quotient correction, then takes the mixed step itself
    (dmrj_twostep, the sweep's own arithmetic).  Returns the per-bond
    relative energy deviations of the two split forms from the FP64 step:
    the precision form's error on identical inputs, apart from the
    trajectory differences that whole sweeps from an unconverged state
    accumulate."""
    import ctypes as C
    name = mixed["precision"]
    rec = []

    def step(n, j, i, iL, iR, psi, mpo_, Lenv, Renv, params):
        out = {}
        for key in ("f64", name, name + "c"):
            h, e = C.c_void_p(), C.c_double()
            ctx.call("dmrj_lanczos2", Lenv[iL].h, Renv[iR].h, psi[iL].h, psi[iR].h, mpo_.device(ctx, iL),
                     mpo_.device(ctx, iR), pkg._lib.PREC[key], C.byref(h), C.byref(e))
            pkg.DeviceTensor(ctx, h).free()
            out[key] = e.value
        rec.append(out)
        pkg.twostep(n, j, i, iL, iR, psi, mpo_, Lenv, Renv, params)

    p = pkg.algvars(st)
    p.maxm, p.minm, p.sweeps, p.cutoff = m, 2, 1, cutoff
    p.precision, p.env_precision, p.correction, p.method = name, "f64", mixed["correction"], "twosite"
    pkg.optmps(st, mpo, p, stepfct=step)
    _free_envs(p)
    res = {"bonds": len(rec)}
    for key in (name, name + "c"):
        r = np.array([abs(x[key] - x["f64"]) / abs(x["f64"]) for x in rec])
        res[key] = {"max": float(r.max()), "median": float(np.median(r)),
                    "bonds_le_1e-10": int(np.sum(r <= 1e-10))}
    log("bond steps of one mixed sweep vs FP64 on identical inputs: %s max %.2e median %.2e; with FP64 "
        "correction max %.2e median %.2e (%d of %d bonds <= 1e-10)"
        % (name, res[name]["max"], res[name]["median"], res[name + "c"]["max"], res[name + "c"]["median"],
           res[name + "c"]["bonds_le_1e-10"], len(rec)))
    return res


def cpu_sweep_baseline(pkg, ctx, psi, mpo, params, budget_s, m, cutoff):
    """The reference's whole two-site bond step (oracle/bondstep.py:
    normalise, simplelanczos with both H_eff*psi, the discarded
    twosite_update, one LAPACK gesdd per centre-charge sector, the split and
    the Lupdate) timed on the host cores at consecutive bonds from the centre
    of the GPU state, left to right, until the budget is spent (at least one
    bond).  A sweep is estimated as 2(L-1) times the mean bond time.  The
    first CPU bond energy is checked against the GPU's FP64 step at the same
    bond from the same state."""
    import multiprocessing
    from oracle import bondstep as bs
    L = len(psi)
    c = L // 2
    j = 1
    while psi.oc < c:
        i = psi.oc
        pkg.twostep(0, j, i, i, i + 1, psi, mpo, params.Lenv, params.Renv, params)
        params.Lenv[i + 1] = pkg.Lupdate(params.Lenv[i], psi[i], mpo, i)
        psi.oc += 1
    nq = psi.nq
    qP = np.asarray(psi.qphys) if nq else np.zeros((mpo[1].shape[1], 0), dtype=np.int64)
    dl = mpo.deltas if nq else [np.zeros((W.shape[0], 0), dtype=np.int64) for W in mpo.H] + \
        [np.zeros((1, 0), dtype=np.int64)]

    def dense(t):
        A, q0, q1 = t.to_dense(nq)
        if not nq:
            q0 = np.zeros((A.shape[0], 0), dtype=np.int64)
            q1 = np.zeros((A.shape[-1], 0), dtype=np.int64)
        return A, q0.astype(np.int64), q1.astype(np.int64)

    Lh, qa, _ = dense(params.Lenv[c])
    A1, ql, qm = dense(psi[c])
    bl, bm = bs.Bond(ql), bs.Bond(qm)
    Lb = bs.lenv_to_blocks(Lh, bl, dl[c - 1])
    AL = bs.mps_to_blocks(A1, bl, qP, bm)
    del Lh, A1
    times, energies, flops = [], [], []
    total = 0.0
    k = 0
    while c + k + 1 <= L:
        i = c + k
        A2, _, qr = dense(psi[i + 1])
        Rh, _, _ = dense(params.Renv[i + 1])
        br = bs.Bond(qr)
        AR = bs.mps_to_blocks(A2, bm, qP, br)
        Rb = bs.renv_to_blocks(Rh, br, dl[i + 1])
        del A2, Rh
        bs.FLOPS[0] = 0.0
        t0 = time.perf_counter()
        E, newL, newR, Lnext, qnew, terr = bs.bond_step_right(Lb, Rb, AL, AR, mpo[i], mpo[i + 1], bl, bm, br, qP,
                                                             dl[i - 1], dl[i], dl[i + 1], m=m, minm=2,
                                                             cutoff=cutoff)
        dt = time.perf_counter() - t0
        times.append(dt)
        energies.append(E)
        flops.append(bs.FLOPS[0])
        total += dt
        log("cpu bond %d|%d: %.2f s, E=%.12f, %.3g flop" % (i, i + 1, dt, E, bs.FLOPS[0]))
        k += 1
        if total > budget_s or k >= 4 or c + k + 1 > L:
            break
        # the centre moves right: new left environment and left sector map
        Lb, AL, bl, bm = Lnext, newR, bs.Bond(qnew), br
    # GPU FP64 drop-in step at the first timed bond from the same state
    gpu_p = pkg.algvars(psi, maxm=m, cutoff=cutoff)
    gpu_p.SvNbond = 0
    pkg.twostep(0, 1, c, c, c + 1, psi, mpo, params.Lenv, params.Renv, gpu_p)
    e_gpu = gpu_p.energy
    per_bond = sum(times) / len(times)
    cores = int(os.environ.get("OMP_NUM_THREADS", multiprocessing.cpu_count()))
    tf = sum(flops) / sum(times) / 1e12
    log("cpu baseline: %d bond steps at %d.., %.2f s per bond, sweep estimate %.0f s, %.4f TFLOP/s; first bond E "
        "%.12f (GPU FP64 step %.12f)" % (len(times), c, per_bond, 2 * (L - 1) * per_bond, tf, energies[0], e_gpu))
    return {"value": round(tf, 5), "unit": "TFLOP/s", "cores": cores, "kind": "port",
            "sweep_s_estimate": round(2 * (L - 1) * per_bond, 1),
            "bond_s": [round(t, 3) for t in times],
            "first_bond_energy": energies[0], "gpu_fp64_first_bond_energy": e_gpu,
            "sample": "%d consecutive two-site bond steps from the centre bond %d|%d of the m=%d state: the "
                      "reference's whole step restated on block-sparse tensors (numpy/OpenBLAS, both H_eff*psi, "
                      "the discarded twosite_update, one gesdd per sector, Lupdate), %.2f s per bond; value = "
                      "algorithmic H_eff*psi + environment TFLOP/s over those steps; sweep estimate = 2(L-1) x "
                      "the mean bond time (bonds within ~6 sites of the chain ends are smaller, so it is an "
                      "upper estimate)" % (len(times), c, c + 1, m, per_bond)}


def bench_measure(pkg, ctx, args, psi, desc, L, m):
    """correlationmatrix(psi, S+, S-) (lib/MPO_MPS/correlationmatrix.jl:21-83) on
    the ramped state: O(L^2) environment updates with one-site operators."""
    Sp, Sm, Sz, O, Id = pkg.spinOps()
    ctx.sync()
    t0 = time.perf_counter()
    M = pkg.correlationmatrix(psi, Sp, Sm)
    ctx.sync()
    dt = time.perf_counter() - t0
    log("correlationmatrix L=%d m=%d: %.3f s, diag mean %.12f, <S+_1 S-_2> %.12f" % (L, m, dt, np.mean(np.diag(M)),
                                                                                  M[0, 1]))
    import tempfile
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "state.npz")
        t1 = time.perf_counter()
        pkg.save_mps(psi, path)
        t2 = time.perf_counter()
        back, _ = pkg.load_mps(path, ctx)
        ctx.sync()
        t3 = time.perf_counter()
        size = os.path.getsize(path)
        back.free()
    log("checkpoint: %.1f MB, save %.3f s, load %.3f s" % (size / 1e6, t2 - t1, t3 - t2))
    out = {"metric": "correlationmatrix(psi, S+, S-) wall time on the GPU-resident MPS", "value": round(dt, 4),
           "unit": "s", "n_gpus": 1, "steps": 1, "warmup": 0, "ms_per_step": round(dt * 1e3, 3),
           "higher_is_better": False, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
           "data": "synthetic (seeded random product state, DMRG ramp)",
           "config": {"workload": "%s, state after the ramp to m=%d, %d environment updates" % (desc, m, L * L),
                      "L": L, "m": m, "diag_mean": float(np.mean(np.diag(M))),
                      "checkpoint": {"bytes": size, "save_s": round(t2 - t1, 4), "load_s": round(t3 - t2, 4)}},
           "roofline": None, "cpu_baseline": None}
    print(json.dumps(out), flush=True)


class _EnvView:
    """An environment list sharing the tensors of another: entries assigned
    through it are new tensors it owns, the shared ones are never freed."""

    def __init__(self, env):
        self.V = list(env.V)
        self.base = {id(x) for x in env.V if x is not None}
        self.new = []

    def __getitem__(self, i):
        return self.V[i]

    def __setitem__(self, i, v):
        self.V[i] = v
        self.new.append(v)

    def __len__(self):
        return len(self.V) - 1

    def free_new(self):
        for t in self.new:
            if t is not None and id(t) not in self.base and t.h:
                t.free()


def centre_step_energies(pkg, ctx, psi, mpo, params):
    """One Lanczos step (dmrj_lanczos2) at the centre bond of the same state
    in every arithmetic: the energy deviation of each split form (plain, and
    with the FP64 Rayleigh-quotient correction) from the FP64 step.  Runs on
    a copy of the state and of the environment lists, so psi and params are
    left as they were."""
    import ctypes as C
    L = len(psi)
    c = L // 2
    j = 1
    st = psi.copy()
    Lv, Rv = _EnvView(params.Lenv), _EnvView(params.Renv)
    p2 = pkg.algvars(st, maxm=params.maxm, cutoff=params.cutoff)
    while st.oc < c:
        i = st.oc
        pkg.twostep(0, j, i, i, i + 1, st, mpo, Lv, Rv, p2)
        Lv[i + 1] = pkg.Lupdate(Lv[i], st[i], mpo, i)
        st.oc += 1
    out = {}
    for name in ("f64", "bf16x3", "bf16x3c", "f16x3", "f16x3c"):
        h, e = C.c_void_p(), C.c_double()
        ctx.call("dmrj_lanczos2", Lv[c].h, Rv[c + 1].h, st[c].h, st[c + 1].h,
                 mpo.device(ctx, c), mpo.device(ctx, c + 1), pkg._lib.PREC[name], C.byref(h), C.byref(e))
        pkg.DeviceTensor(ctx, h).free()
        out[name] = e.value
    Lv.free_new()
    Rv.free_new()
    st.free()
    ref = out["f64"]
    res = {k: {"energy": v, "rel_dE": abs(v - ref) / abs(ref)} for k, v in out.items()}
    log("centre bond %d|%d Lanczos step: " % (c, c + 1) +
        ", ".join("%s %.3e" % (k, v["rel_dE"]) for k, v in res.items() if k != "f64"))
    return {"bond": [c, c + 1], "variants": res}


def bench_compare(pkg, ctx, args, psi, mpo, desc, L, m, base_params):
    """Variants of the same sweeps from copies of one state: the plain FP64
    drop-in (SVD backend 0) is the reference for time and energy.  Energies
    after a few sweeps differ by the variants' different truncation paths
    unless the state is converged; the centre-bond Lanczos step from one
    state isolates the arithmetic."""
    variants = [("f64", None, 0), ("f64", None, 5), ("bf16x3", "f64", 5), ("f16x3", "f64", 5)]
    # the centre-bond step (on a copy of the ramp state) needs the ramp's
    # environments; free them afterwards so that only one variant's
    # environments are resident at a time; every variant starts from the
    # unmodified ramp state
    centre = centre_step_energies(pkg, ctx, psi, mpo, base_params)
    for env in (base_params.Lenv, base_params.Renv):
        for e in (env.V if env is not None else []):
            if e is not None:
                e.free()
    base_params.Lenv = base_params.Renv = None
    rows = []
    for prec, envp, backend in variants:
        st = psi.copy()
        ctx.set_svd_backend(backend)
        params = pkg.algvars(st)
        ctx.sync()
        t0 = time.perf_counter()
        pkg.dmrg(st, mpo, m=m, sweeps=args.steps, cutoff=args.cutoff, params=params, precision=prec,
                 env_precision=envp)
        ctx.sync()
        elapsed = time.perf_counter() - t0
        hist = params.history[-args.steps:]
        rows.append({"precision": prec, "env_precision": envp or prec, "svd_backend": backend,
                     "sweep_s": [round(h["time"], 4) for h in hist], "energies": [h["energy"] for h in hist],
                     "storeD": hist[-1]["storeD"], "maxtrunc": hist[-1]["maxtrunc"],
                     "biggestm": hist[-1]["biggestm"], "wall_s": elapsed})
        log("variant %-6s env %-6s svd %d: sweeps %s s, E %s" % (prec, envp or prec, backend, rows[-1]["sweep_s"],
                                                                 ["%.12f" % e for e in rows[-1]["energies"]]))
        st.free()
        for env in (params.Lenv, params.Renv):
            for e in (env.V if env is not None else []):
                if e is not None:
                    e.free()
    ctx.set_svd_backend(0)
    ref = rows[0]
    t_ref = sum(ref["sweep_s"]) / len(ref["sweep_s"])
    for r in rows:
        t = sum(r["sweep_s"]) / len(r["sweep_s"])
        r["speedup_vs_fp64_dropin"] = round(t_ref / t, 3)
        r["max_rel_dE"] = max(abs(a - b) / abs(b) for a, b in zip(r["energies"], ref["energies"]))
        n = min(len(r["storeD"]), len(ref["storeD"]))
        r["max_abs_d_spectrum"] = max([abs(a - b) for a, b in zip(r["storeD"][:n], ref["storeD"][:n])] or [0.0])
    best = rows[2]
    out = {
        "metric": "DMRG sweep time, mixed vs FP64 GPU from the same state, and |dE/E| against the FP64 drop-in",
        "value": best["speedup_vs_fp64_dropin"], "unit": "x (sweep-time ratio, BF16x3 + FP64 env + Gram SVD)",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(t_ref * 1e3, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64 / bf16x3 / f16x3",
        "data": "synthetic (seeded random product state)",
        "config": {"workload": "%s, two-site DMRG sweeps at m=%d from one state, Krylov 2, cutoff %g"
                               % (desc, m, args.cutoff), "L": L, "m": m},
        "variants": rows, "centre_step": centre,
        "roofline": None, "cpu_baseline": None,
    }
    print(json.dumps(out), flush=True)


def bench_tdvp(pkg, ctx, args):
    """Real-time two-site TDVP (algs/time.jl:288-362): Heisenberg chain from
    the Neel state, U(1), prefactor -i dt per half-sweep, Krylov dimension
    --krylov, cutoff 1e-12, bond dimension capped at m.  --warmup half-sweeps
    untimed, then --steps timed half-sweeps."""
    L = args.L or 100
    m = args.m or 1024
    qphys = np.array([[1], [-1]])
    mpo = pkg.MPO(pkg.makeMPO(pkg.heisenbergMPO(), 2, L).H, qphys=qphys)
    psi = pkg.product_state([i % 2 for i in range(L)], 2, qphys=qphys, ctx=ctx)
    Le = Re = None
    for _ in range(args.warmup):
        _, Le, Re = pkg.tdvp_twosite(psi, mpo, m=m, cutoff=1e-12, r=args.krylov, prefactor=-1j * args.dt,
                                     Lenv=Le, Renv=Re)
    # timed region without per-kernel events (they would add a record pair to
    # every one of these small launches)
    ctx.sync()
    ctx.flops(reset=True)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        _, Le, Re = pkg.tdvp_twosite(psi, mpo, m=m, cutoff=1e-12, r=args.krylov, prefactor=-1j * args.dt,
                                     Lenv=Le, Renv=Re)
    ctx.sync()
    elapsed = time.perf_counter() - t0
    fl_heff, fl_env = ctx.flops()
    # kernel statistics from two further, profiled half-sweeps
    ctx.profile(True)
    for k in ("gemm_f64", "mix", "copy", "svd"):
        ctx.kernel_stats(k, reset=True)
    for _ in range(2):
        _, Le, Re = pkg.tdvp_twosite(psi, mpo, m=m, cutoff=1e-12, r=args.krylov, prefactor=-1j * args.dt,
                                     Lenv=Le, Renv=Re)
    ctx.sync()
    stats = {k: ctx.kernel_stats(k) for k in ("gemm_f64", "mix", "copy", "svd")}
    ctx.profile(False)
    chi = max(psi.bond_dims())
    sz = pkg.local_expect(psi, np.diag([0.5, -0.5]), L // 2)
    g = stats["gemm_f64"]
    g_tflops = g["flops"] / (g["ms"] * 1e-3) / 1e12 if g["ms"] > 0 else 0.0
    log("tdvp: %d half-sweeps after %d (+2 profiled), t=%.3f, max bond %d, <Sz_%d>=%.12f, %.3f s per "
        "half-sweep, gemm %.2f TFLOP/s, svd %.3f s per profiled half-sweep"
        % (args.steps, args.warmup, args.dt * (args.warmup + args.steps + 2), chi, L // 2, sz, elapsed / args.steps,
           g_tflops, stats["svd"]["ms"] / 2e3))
    # at the bond dimensions this workload reaches the half-sweep is bound by
    # latency, not flops: the throughput is bond steps (two Krylov
    # exponentials, an SVD and an environment update) per second
    out = {
        "metric": "real-time two-site TDVP half-sweep (ComplexF64, U(1)): bond steps per second",
        "value": round((L - 1) * args.steps / elapsed, 2), "unit": "bond steps/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(elapsed / args.steps * 1e3, 3), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "c128 (split planes, FP64 arithmetic)", "data": "synthetic (Neel state)",
        "config": {"workload": "Heisenberg L=%d Neel start, real-time TDVP dt=%g Krylov %d, m<=%d, cutoff 1e-12"
                               % (L, args.dt, args.krylov, m), "max_bond": chi, "sz_centre": sz,
                   "contraction_tflops": (fl_heff + fl_env) / elapsed / 1e12},
        "graph": ctx.graph_stats() if args.graph else None,
        "roofline": {"bound": "mfma", "kernel": "k_gemm_f64 (complexified plans)", "achieved": round(g_tflops, 3),
                     "peak": PEAK_FP64, "unit": "TFLOP/s", "frac": round(g_tflops / PEAK_FP64, 4), "traffic": None,
                     "note": "launch durations from two profiled half-sweeps right after the timed region (events "
                             "on every launch would slow the timed half-sweeps by ~40 %); at bond dimension 4 "
                             "the launches are latency-bound"},
        "cpu_baseline": None,
    }
    print(json.dumps(out), flush=True)


def cpu_baseline(pkg, psi, mpo, params, budget_s, method="twosite"):
    """Time the oracle's block-sparse H_eff*psi at the centre bond of the
    current (m = 1024) state on the host cores (the single-site form at the
    centre site for 3S).  The GPU state is first moved to the centre (untimed
    bond steps) so that both environments are current."""
    import multiprocessing
    from oracle import blocks as ob
    L = len(psi)
    c = L // 2
    j = 1
    while psi.oc < c:
        i = psi.oc
        pkg.twostep(0, j, i, i, i + 1, psi, mpo, params.Lenv, params.Renv, params)
        params.Lenv[i + 1] = pkg.Lupdate(params.Lenv[i], psi[i], mpo, i)
        psi.oc += 1
    nq = psi.nq
    Lh, qa, _ = params.Lenv[c].to_dense(nq)
    Rh, qb, _ = params.Renv[c + 1].to_dense(nq)
    A1, ql, _ = psi[c].to_dense(nq)
    A2, _, qr = psi[c + 1].to_dense(nq)
    theta = np.tensordot(A1, A2, axes=(2, 0))
    plan = ob.Heff2Plan(Lh, Rh, mpo.H[c - 1], mpo.H[c], ql, psi.qphys, qr, mpo.deltas[c - 1],
                        mpo.deltas[c + 1])
    if method == "3S":
        Rc, _, _ = params.Renv[c].to_dense(nq)
        _, _, qc = psi[c].to_dense(nq)
        theta = A1
        plan = ob.Heff1Plan(Lh, Rc, mpo.H[c - 1], ql, psi.qphys, qc, mpo.deltas[c - 1], mpo.deltas[c])
    n = 0
    t0 = time.perf_counter()
    while True:
        plan.apply(theta)
        n += 1
        if time.perf_counter() - t0 > budget_s or n >= 50:
            break
    dt = (time.perf_counter() - t0) / n
    tf = plan.flops / dt / 1e12
    cores = int(os.environ.get("OMP_NUM_THREADS", multiprocessing.cpu_count()))
    form = "single-site H_eff*A at the centre site %d" % c if method == "3S" else \
        "two-site H_eff*psi at the centre bond %d|%d" % (c, c + 1)
    log("cpu baseline: %d %s applications (chi %d x %d), %.3f s each, %.4f TFLOP/s"
        % (n, form, theta.shape[0], theta.shape[-1], dt, tf))
    return {"value": round(tf, 5), "unit": "TFLOP/s", "cores": cores, "kind": "port",
            "sample": "%d block-sparse %s applications (chi %dx%d) of the m=%d state, numpy/OpenBLAS oracle, "
                      "%.3f s each" % (n, form, theta.shape[0], theta.shape[-1], params.maxm, dt)}


if __name__ == "__main__":
    main()
