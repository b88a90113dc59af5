"""Benchmark: leverage probe solves per second on one MI355X.

Workload: the 2,000,000-worker x T=5 AKM panel with 100,000 firms (Pareto 1.1
firm sizes truncated at 5,000, mobility 0.1), leave-out connected set, match
deletion; Counter-V1 Rademacher leverage probes solved by CMG-preconditioned
PCG on the degree-4 hybrid Laplacian to relative residual 1e-8.

A step is one batch of `--width` probes through the whole per-probe path on
the GPU: atoms -> exact right-hand sides -> Schur right-hand side ->
lockstep certified PCG with the CMG V-cycle -> extraction and the complete
residual gate -> leverage moments.  The headline runs the learned smoother
weights (the library's built-in trained network, kept only when its
calibration batch needs fewer iterations) with batched probes and the FP32
V-cycle.  Inputs (panel, hierarchy) are resident in HBM before the timed
region.  Reported beside it: the same batches with the plain CMG cycle, the
plain single-RHS FP64 path on the same GPU, the full KSS estimate (fit, 256
leverage and 256 target probes) wall time learned+batched against plain
single-RHS FP64 with the errors between them, and the oracle (the CPU
restatement of the reference engine) on the host's cores on a bounded
sample.

The roofline entry is the dominant kernel of the timed region, the finest
direction and A p of the Krylov iteration (k_segments<double, 64, 1,
OpDirectionCurvature<64, float>>, pass 1), timed with HIP events on the
engine stream; `traffic` comes from the
committed rocprofv3 FETCH_SIZE / WRITE_SIZE summary of that kernel for this
configuration when present.  The finest prolong + smooth pass of the V-cycle
(the V-cycle SpMV) is reported beside it the same way.

Prints one JSON line on stdout; progress goes to stderr.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def log(msg):
    print(f"[bench {time.strftime('%H:%M:%S')}] {msg}", file=sys.stderr, flush=True)


def measured_traffic(key, name="ap_traffic.json"):
    """HBM bytes per launch of a kernel from a committed counter summary."""
    path = os.path.join(ROOT, "profiles", name)
    try:
        with open(path) as fh:
            data = json.load(fh)
    except (OSError, ValueError):
        return None
    entry = data.get(key)
    return None if entry is None else entry.get("hbm_bytes_per_launch")


def host_description():
    """CPU model, nproc and the cores this process may run on."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    usable = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    return {"cpu_model": model, "nproc": os.cpu_count() or 1, "usable_cores": usable}


def cpu_threads(requested):
    """Oracle threads: every core this process may use, capped at the 16-core
    host share of one GPU, unless given explicitly."""
    if requested > 0:
        return requested
    return max(1, min(16, host_description()["usable_cores"]))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=8)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--width", type=int, default=64, help="probes per batch (1..64)")
    ap.add_argument("--precision", type=int, default=32, choices=(32, 64),
                    help="V-cycle precision (Krylov recurrences are always FP64)")
    ap.add_argument("--tolerance", type=float, default=1e-8)
    ap.add_argument("--learned", type=int, default=1,
                    help="use the built-in learned smoother weights for the headline (0: plain CMG)")
    ap.add_argument("--autotune", type=int, default=1,
                    help="pick the batch width and segment chunk size by a timed search first "
                         "(untimed; --width is then ignored)")
    ap.add_argument("--plain-steps", type=int, default=16,
                    help="single-RHS FP64 solves timed for the comparison (0 skips)")
    ap.add_argument("--kss-probes", type=int, default=256, help="full KSS estimate probes (0 skips)")
    ap.add_argument("--subsamples", type=int, default=64,
                    help="uneven-size subsamples of the sample-batch estimate (0 skips)")
    ap.add_argument("--subsample-streams", default="1,4,8,16",
                    help="stream counts of the concurrent sample-batch runs beside lockstep ('' for lockstep only)")
    ap.add_argument("--controls", type=int, default=10,
                    help="year-dummy controls of the controls estimate on the same panel (0 skips)")
    ap.add_argument("--compare-plain", type=int, default=1,
                    help="also run the KSS estimate on the plain single-RHS FP64 path and report the error")
    ap.add_argument("--exact", type=int, default=1, help="time the dense exact estimator on the small panel")
    ap.add_argument("--observation", type=int, default=1,
                    help="KSS estimate under observation deletion on one 40k-worker subsample (0 skips)")
    ap.add_argument("--learned-eval", type=int, default=2,
                    help="leverage batches of the held-out 10M-worker smoother evaluation (0 skips)")
    ap.add_argument("--cpu-solves", type=int, default=64, help="oracle probe solves per timed run (0 skips)")
    ap.add_argument("--cpu-runs", type=int, default=5, help="warm oracle runs; the median is reported")
    ap.add_argument("--cpu-threads", type=int, default=0,
                    help="oracle threads (0: every core this process may use, at most 16 per GPU)")
    ap.add_argument("--scale", type=float, default=1.0, help="panel size factor (1.0 = full size)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported "
              "(the engine runs one process on one MI355X).", file=sys.stderr)
        sys.exit(2)

    import fevc_amd as F
    from fevc_amd import synth

    spec = synth.SCALE if args.scale == 1.0 else synth.scaled(synth.SCALE, args.scale)
    log(f"generate {spec.name}")
    data = synth.generate(spec)
    cols = [data[k] for k in ("worker", "firm", "deletion", "outcome", "frequency", "target")]
    year_z = year_y = None
    if args.controls > 0:
        # the same panel with `controls` year dummies and a year effect in the outcome
        year_z, year_y = synth.year_controls(spec, data, years=max(args.controls, spec.periods) + 1)
        year_z = year_z[:, :args.controls]
    del data
    t0 = time.perf_counter()
    eng = F.Engine(*cols)
    prep_s = time.perf_counter() - t0
    d = eng.dims
    log(f"prepared in {prep_s:.1f}s: {d}")

    copy_bytes = 1 << 29
    eng.stream_copy(copy_bytes)  # first call allocates and faults the pages in
    copy_ms = eng.stream_copy(copy_bytes)
    log(f"stream copy {copy_bytes >> 20} MiB: {2 * copy_bytes / copy_ms / 1e6:.0f} GB/s")

    smoother = None
    eng.configure(args.precision, min(args.width, 64), min(max(args.width // 2, 1), 32))
    if args.learned:
        t0 = time.perf_counter()
        eng.smoother(1)
        smoother = eng.smoother_status()
        smoother["setup_seconds"] = round(time.perf_counter() - t0, 3)
        log(f"tuned preconditioner: calibration iterations (reference, learned, shaped, both) "
            f"{smoother['calibration']}; learned weights {smoother['active']}, cycle levels {smoother['shape']}")
    tuning = None
    if args.autotune:
        t0 = time.perf_counter()
        tuning = eng.autotune(args.precision, args.tolerance, batches=1)
        args.width = tuning["width"]
        log(f"autotune ({time.perf_counter() - t0:.1f}s): width {tuning['width']}, chunk "
            f"{tuning['chunk']}; trials " + ", ".join(f"{int(w)}/{int(c)}: {r:.0f}"
                                                      for w, c, r in tuning["trials"]))
        tuning = {"width": tuning["width"], "chunk": tuning["chunk"],
                  "trials_solves_per_s": [[int(w), int(c), round(r, 1)] for w, c, r in tuning["trials"]]}
    # warmup also builds the device copy of the problem and hierarchy
    log(f"warmup: {args.warmup} batch(es) of {args.width}, fp{args.precision} V-cycle")
    eng.time_leverage(args.width, args.precision, args.tolerance, max(args.warmup, 1))
    log(f"timed: {args.steps} batch(es)")
    c0 = eng.counters()
    t0 = time.perf_counter()
    tm = eng.time_leverage(args.width, args.precision, args.tolerance, args.steps)
    wall = time.perf_counter() - t0
    c1 = eng.counters()
    refinement = {"rounds": c1["refine_rounds"] - c0["refine_rounds"],
                  "refined_columns": c1["refined_columns"] - c0["refined_columns"],
                  "batches": c1["batches"] - c0["batches"]}
    solves = tm["solves"]
    value = solves / wall
    iters = tm["iterations"] / max(solves, 1)
    kernel_s = tm["kernel_ms"] / 1e3
    achieved = tm["kernel_bytes"] / kernel_s / 1e9 if kernel_s > 0 else 0.0
    ps = eng.kernel_timing(1)
    ps_gbs = ps["bytes"] / (ps["ms"] / 1e3) / 1e9 if ps["ms"] > 0 else 0.0
    log(f"batched: {value:.1f} solves/s, {iters:.1f} it/solve, refinement {refinement}; finest A p "
        f"{tm['kernel_ms'] * 1e3:.1f} us x {tm['kernel_launches']} -> {achieved:.0f} GB/s; finest prolong-smooth "
        f"{ps['ms'] * 1e3:.1f} us x {ps['launches']} -> {ps_gbs:.0f} GB/s")

    plain_batched = None
    if args.learned and args.plain_steps > 0:
        eng.smoother(0)
        eng.time_leverage(args.width, args.precision, args.tolerance, 1)
        t0 = time.perf_counter()
        tb = eng.time_leverage(args.width, args.precision, args.tolerance, args.steps)
        bw = time.perf_counter() - t0
        plain_batched = {"solves_per_s": round(tb["solves"] / bw, 3),
                         "iterations_per_solve": round(tb["iterations"] / max(tb["solves"], 1), 2)}
        log(f"plain CMG, same batches: {plain_batched['solves_per_s']:.1f} solves/s, "
            f"{plain_batched['iterations_per_solve']:.1f} it/solve")
    eng.smoother(0)
    plain = None
    if args.plain_steps > 0:
        eng.time_leverage(1, 64, args.tolerance, 1)
        t0 = time.perf_counter()
        tp = eng.time_leverage(1, 64, args.tolerance, args.plain_steps)
        pw = time.perf_counter() - t0
        plain = {"solves_per_s": round(tp["solves"] / pw, 3), "solves": tp["solves"],
                 "iterations_per_solve": round(tp["iterations"] / max(tp["solves"], 1), 2)}
        log(f"plain single-RHS FP64: {plain['solves_per_s']:.1f} solves/s")

    kss = None
    if args.kss_probes > 0:
        eng.configure(args.precision, min(args.width, 64), min(max(args.width // 2, 1), 32))
        eng.smoother(1 if args.learned else 0)
        t0 = time.perf_counter()
        res = eng.solve(probes=args.kss_probes, tolerance=args.tolerance, fit_tolerance=args.tolerance)
        kw = time.perf_counter() - t0
        arr = eng.arrays(args.kss_probes)
        kss = {"wall_seconds": round(kw, 3), "probes": args.kss_probes,
               "corrected": res["corrected"], "numerical_mcse": res["numerical_mcse"]}
        log(f"KSS estimate ({args.kss_probes} probes): {kw:.2f}s, corrected {res['corrected']}")
        # the command's default adds the numerical MCSE (replayed leverage probes)
        t0 = time.perf_counter()
        _, num = eng.solve_numerical(probes=args.kss_probes, tolerance=args.tolerance,
                                     fit_tolerance=args.tolerance)
        nw = time.perf_counter() - t0
        kss["numerical_mcse_wall_seconds"] = round(nw, 3)
        kss["numerical_mcse"] = [float(v) for v in num["mcse"]]
        kss["numerical_mcse_status"] = int(num["status"])
        log(f"KSS estimate with numerical MCSE replay: {nw:.2f}s, status {num['status']}, mcse {list(num['mcse'])}")
        if args.compare_plain:
            # the same estimate on the plain single-RHS FP64 path, identical draws
            eng.configure(64, 1, 1)
            eng.smoother(0)
            t0 = time.perf_counter()
            ref = eng.solve(probes=args.kss_probes, tolerance=args.tolerance, fit_tolerance=args.tolerance)
            pw = time.perf_counter() - t0
            ref_arr = eng.arrays(args.kss_probes)
            rel = {k: float(np.max(np.abs(np.asarray(res[k]) - np.asarray(ref[k])) /
                                   np.maximum(np.abs(np.asarray(ref[k])), 1e-300)))
                   for k in ("plugin", "correction", "corrected")}
            lev = {k: float(np.max(np.abs(arr[k] - ref_arr[k])))
                   for k in ("projection_share", "residual_share")}
            kss["plain_fp64_wall_seconds"] = round(pw, 3)
            kss["wall_speedup_vs_plain_fp64"] = round(pw / kw, 2)
            kss["max_rel_error_vs_plain_fp64"] = rel
            kss["max_abs_leverage_error_vs_plain_fp64"] = lev
            log(f"plain FP64 KSS: {pw:.2f}s; component rel. error {rel}; leverage abs. error {lev}")
    eng.release()

    controls = None
    if year_z is not None:
        # the same graph with year-dummy controls: the joint nuisance (the
        # command's default: Q projection solves, the joint fit, 3R probe
        # solves) and the fixed offset (one more working fit)
        ccols = list(cols)
        ccols[3] = year_y
        probes = args.kss_probes or 256
        t0 = time.perf_counter()
        ce = F.Engine(*ccols, controls=year_z)
        cprep = time.perf_counter() - t0
        try:
            ce.configure(args.precision, min(args.width, 64), min(max(args.width // 2, 1), 32))
            ce.smoother(1 if args.learned else 0)
            # device set-up (a few-probe estimate can meet a block whose
            # estimated joint leverage reaches one, so the warm-up uses all)
            ce.solve(probes=probes, tolerance=args.tolerance, fit_tolerance=args.tolerance)
            q = year_z.shape[1]
            controls = {"controls": q, "kind": "year dummies", "prepare_seconds": round(cprep, 1),
                        "probes": probes}
            for nuisance in ("joint", "fixed_offset"):
                t0 = time.perf_counter()
                r = ce.solve(probes=probes, tolerance=args.tolerance, fit_tolerance=args.tolerance,
                             nuisance=nuisance)
                w = time.perf_counter() - t0
                cf = ce.control_fit()
                rhs = q + 1 + (nuisance == "fixed_offset") + 3 * probes
                controls[nuisance] = {"wall_seconds": round(w, 3), "rhs_solves": rhs,
                                      "rhs_solves_per_s": round(rhs / w, 1), "corrected": r["corrected"],
                                      "control_rcond": cf["rcond"],
                                      "max_abs_year_coefficient": float(np.max(np.abs(cf["beta"])))}
                log(f"controls ({q} year dummies, {nuisance}): KSS {w:.2f}s, {rhs / w:.0f} RHS solves/s, "
                    f"corrected {r['corrected']}")
                if nuisance == "joint":
                    joint_arrays = ce.arrays(probes)
                    joint_result = r
            if kss:
                controls["joint_wall_vs_no_controls"] = round(controls["joint"]["wall_seconds"] /
                                                              kss["wall_seconds"], 3)
            if args.compare_plain:
                # the joint estimate on the plain single-RHS FP64 path, identical draws:
                # per-component and per-leverage error of the batched mixed-precision run
                ce.configure(64, 1, 1)
                ce.smoother(0)
                t0 = time.perf_counter()
                ref = ce.solve(probes=probes, tolerance=args.tolerance, fit_tolerance=args.tolerance)
                pw = time.perf_counter() - t0
                ref_arr = ce.arrays(probes)
                controls["joint"]["plain_fp64_wall_seconds"] = round(pw, 3)
                controls["joint"]["max_rel_error_vs_plain_fp64"] = {
                    k: float(np.max(np.abs(np.asarray(joint_result[k]) - np.asarray(ref[k])) /
                                    np.maximum(np.abs(np.asarray(ref[k])), 1e-300)))
                    for k in ("plugin", "correction", "corrected")}
                controls["joint"]["max_abs_leverage_error_vs_plain_fp64"] = {
                    k: float(np.max(np.abs(joint_arrays[k] - ref_arr[k])))
                    for k in ("projection_share", "residual_share")}
                log(f"controls, plain FP64 joint KSS: {pw:.2f}s; errors "
                    f"{controls['joint']['max_rel_error_vs_plain_fp64']}")
        finally:
            ce.release()

    batch = None
    if args.subsamples > 0:
        # many independent region-year subsamples of uneven size (log-normal
        # workers, median 40 000, clipped to 5 000-500 000; firms = workers /
        # 20; T = 4; mobility U(0.05, 0.25)), observation deletion, 128 probes
        # each: in lockstep (every PCG pass one launch over all 64 samples) and
        # concurrently on 4, 8 and 16 streams, against one sample at a time
        sc, sspecs = synth.subsamples(count=args.subsamples, scale=args.scale)
        t0 = time.perf_counter()
        sb = F.SampleBatch(sc["sample"], *[sc[k] for k in ("worker", "firm", "deletion", "outcome", "frequency",
                                                             "target")],
                           deletion_mode="observation", threads=cpu_threads(0))
        bprep = time.perf_counter() - t0
        try:
            n = int(sb.samples)
            stream_counts = [int(x) for x in args.subsample_streams.split(",") if x.strip()]
            # the first solve: device buffers and every sample's smoother
            # calibration (learned weights and cycle shape kept where they help)
            t0 = time.perf_counter()
            sb.solve(probes=128, tolerance=args.tolerance, fit_tolerance=args.tolerance,
                     precision=args.precision, streams=n, lockstep=True)
            first = time.perf_counter() - t0
            for lockstep in ([False] if stream_counts else []) + [True]:  # warm: buffers, code objects
                sb.solve(probes=16, tolerance=args.tolerance, fit_tolerance=args.tolerance,
                         precision=args.precision, streams=n if lockstep else 4, lockstep=lockstep)
            runs = {}
            for streams in stream_counts:
                t0 = time.perf_counter()
                res = sb.solve(probes=128, tolerance=args.tolerance, fit_tolerance=args.tolerance,
                               precision=args.precision, streams=streams)
                runs[streams] = (time.perf_counter() - t0, res)
            lock = []
            for _ in range(3):
                t0 = time.perf_counter()
                res = sb.solve(probes=128, tolerance=args.tolerance, fit_tolerance=args.tolerance,
                               precision=args.precision, streams=n, lockstep=True)
                lock.append((time.perf_counter() - t0, res))
            wl, rl = min(lock, key=lambda x: x[0])
            w1, r1 = runs.get(1, (float("nan"), rl))
            concurrent = [s for s in runs if s > 1]
            best = min(concurrent, key=lambda s: runs[s][0]) if concurrent else None
            ws = runs[best][0] if best else float("nan")
            same = all(np.array_equal(a[k], b[k]) for other in [runs[s][1] for s in runs] + [r for _, r in lock]
                       for a, b in zip(r1, other) for k in ("plugin", "correction", "corrected", "mcse"))
            rows = [int(r["rows"]) for r in rl]
            batch = {"subsamples": len(rl), "deletion": "observation", "probes": 128,
                     "workers_median": int(np.median([s.workers for s in sspecs])),
                     "workers_range": [min(s.workers for s in sspecs), max(s.workers for s in sspecs)],
                     "stored_rows_total": int(sum(rows)), "prepare_seconds": round(bprep, 2),
                     "lockstep_seconds": round(wl, 3),
                     "lockstep_first_solve_seconds": round(first, 3),
                     "lockstep_runs_seconds": [round(w, 3) for w, _ in lock],
                     "one_at_a_time_seconds": round(w1, 3) if 1 in runs else None,
                     "concurrent_seconds": {str(s): round(runs[s][0], 3) for s in concurrent},
                     "best_streams": best,
                     "speedup_lockstep_vs_one_at_a_time": round(w1 / wl, 2) if 1 in runs else None,
                     "speedup_lockstep_vs_best_streams": round(ws / wl, 2) if best else None,
                     "probe_solves_per_s": round(3 * 128 * len(rl) / wl, 1),
                     "results_identical": bool(same)}
            log(f"subsample batch ({len(rl)} samples, {sum(rows)} rows): lockstep {wl:.3f}s "
                f"(first solve with the smoother calibration {first:.3f}s), "
                + ", ".join(f"{s} streams {runs[s][0]:.2f}s" for s in runs) + f", identical {same}")
        finally:
            sb.release()

    exact = None
    if args.exact:
        # dense exact estimator on the small plain-method panel (2000 workers x
        # 5 periods, 200 firms, mobility 0.2, seed 1): exact leverages, EXACT
        # algorithm with the dimension limit raised above its W + F - 1
        small = synth.PanelSpec(workers=2000, firms=200, periods=5, mobility=0.2, seed=1,
                                name="akm-2k-workers-200-firms-T5")
        sd = synth.generate(small)
        se = F.Engine(*[np.asarray(sd[k]) for k in ("worker", "firm", "deletion", "outcome", "frequency",
                                                    "target")])
        try:
            se.solve_exact(exact_limit=10000)  # warm (device set-up, code objects)
            t0 = time.perf_counter()
            res, rec = se.solve_exact(exact_limit=10000)
            ew = time.perf_counter() - t0
        finally:
            se.release()
        exact = {"panel": small.name, "parameters": rec["parameters"], "deletion_units": rec["deletion_units"],
                 "wall_seconds": round(ew, 4), "factor_fit_seconds": round(rec["seconds_factor"], 4),
                 "cell_pass_seconds": round(rec["seconds_cells"], 4),
                 "max_leverage": rec["max_leverage"], "information_rcond_lower_bound": rec["information_rcond"],
                 "corrected": res["corrected"]}
        log(f"exact estimator ({small.name}, {rec['parameters']} parameters): {ew * 1e3:.1f} ms, "
            f"corrected {res['corrected']}")

    learned_eval = None
    if args.learned_eval > 0:
        # held-out evaluation of the built-in smoother weights: 10M workers x
        # 3 periods, 500k firms with Pareto(1.05) sizes, mobility 0.1, 256
        # probes, reference cycle against the tuned preconditioner
        big = synth.PanelSpec(workers=10_000_000, firms=500_000, periods=3, mobility=0.1, pareto_shape=1.05,
                              size_cap=5000.0, seed=7, name="akm-10M-workers-500k-firms-T3")
        log(f"learned smoother evaluation: generate {big.name}")
        bd = synth.generate(big)
        bcols = [bd[k] for k in ("worker", "firm", "deletion", "outcome", "frequency", "target")]
        del bd
        t0 = time.perf_counter()
        be = F.Engine(*bcols)
        bprep = time.perf_counter() - t0
        del bcols
        log(f"learned smoother evaluation: prepared in {bprep:.1f}s: {be.dims}")
        try:
            be.configure(args.precision, 64, 32)
            learned_eval = {"panel": big.name, "dims": be.dims, "prepare_seconds": round(bprep, 1),
                            "leverage_batches": args.learned_eval, "probes": 256,
                            "network": "8 features, input layer + 3 message-passing layers of width 64, "
                                       "trained on 200 panels of 50k-1M workers",
                            "inference": "BF16 weights and activations, FP32 accumulation, "
                                         "layer products on the matrix cores"}
            for mode, name in ((0, "plain"), (1, "learned")):
                be.smoother(mode)
                entry = {}
                if mode:
                    st = be.smoother_status()
                    entry["calibration"] = st["calibration"]
                    entry["shape"] = st["shape"]
                    entry["active"] = st["active"]
                be.time_leverage(64, args.precision, args.tolerance, 1)
                t0 = time.perf_counter()
                tl = be.time_leverage(64, args.precision, args.tolerance, args.learned_eval)
                lw = time.perf_counter() - t0
                entry["solves_per_s"] = round(tl["solves"] / lw, 1)
                entry["iterations_per_solve"] = round(tl["iterations"] / max(tl["solves"], 1), 2)
                t0 = time.perf_counter()
                res = be.solve(probes=256, tolerance=args.tolerance, fit_tolerance=args.tolerance)
                entry["kss_wall_seconds"] = round(time.perf_counter() - t0, 3)
                entry["corrected"] = res["corrected"]
                learned_eval[name] = entry
                log(f"learned smoother evaluation, {name}: {entry['solves_per_s']:.0f} solves/s, "
                    f"{entry['iterations_per_solve']:.1f} it/solve, KSS {entry['kss_wall_seconds']:.2f}s")
            pl, le = learned_eval["plain"], learned_eval["learned"]
            learned_eval["solve_speedup"] = round(le["solves_per_s"] / pl["solves_per_s"], 3)
            learned_eval["kss_speedup"] = round(pl["kss_wall_seconds"] / le["kss_wall_seconds"], 3)
            learned_eval["max_rel_component_difference"] = max(
                abs(a - b) / max(abs(b), 1e-300) for a, b in zip(le["corrected"], pl["corrected"]))
        finally:
            be.release()

    cpu = None
    threads = cpu_threads(args.cpu_threads)
    if args.cpu_solves > 0:
        from oracle import oracle as O

        log("cpu baseline: oracle prepare")
        p = O.Problem(*cols)
        host = host_description()
        log(f"cpu baseline: {host['cpu_model']}, nproc {host['nproc']}, usable {host['usable_cores']}; "
            f"1 warm-up + {args.cpu_runs} timed runs of {args.cpu_solves} probe solves on {threads} threads")
        p.time_probe_solves(args.cpu_solves, first=0, threads=threads, probe_tolerance=args.tolerance)
        runs = []
        for r in range(max(1, args.cpu_runs)):
            sec, it = p.time_probe_solves(args.cpu_solves, first=(r + 1) * args.cpu_solves, threads=threads,
                                          probe_tolerance=args.tolerance)
            runs.append((sec, it))
            log(f"cpu baseline run {r + 1}: {args.cpu_solves / sec:.2f} solves/s")
        secs = sorted(sec for sec, _ in runs)
        med = secs[len(secs) // 2]
        its = sum(it for _, it in runs) / (len(runs) * args.cpu_solves)
        cpu = {"value": round(args.cpu_solves / med, 3), "unit": "probe solves/s", "cores": threads,
               "kind": "port",
               "sample": f"median of {len(runs)} warm runs (after 1 warm-up) of {args.cpu_solves} leverage probe "
                         f"solves each (atoms, rhs, CMG-PCG to {args.tolerance:g}, extraction) of the same panel; "
                         f"median run {med:.2f}s, {its:.1f} it/solve",
               "runs_seconds": [round(s_, 3) for s_ in secs], **host}
        log(f"cpu baseline: {cpu['value']:.2f} solves/s (median)")

    traffic = measured_traffic(f"w{args.width}_fp{args.precision}")
    observation = None
    if args.observation:
        # one subsample of the many-subsample observation-deletion workload at its
        # median size: 40 000 workers x 4 periods, 2 000 firms (W / 20), mobility
        # 0.15, Pareto firm sizes; 128 probes, probe solves to rel.res 1e-8
        sub = synth.PanelSpec(workers=40000, firms=2000, periods=4, mobility=0.15, seed=1,
                              name="akm-40k-workers-2k-firms-T4-observation")
        od = synth.generate(sub)
        ocols = [np.asarray(od[k]) for k in ("worker", "firm", "deletion", "outcome", "frequency", "target")]
        observation = {"panel": sub.name, "probes": 128}
        for mode in ("observation", "match"):
            oe = F.Engine(*ocols, deletion_mode=mode)
            try:
                oe.configure(args.precision, 64, 32)
                oe.solve(probes=16, tolerance=args.tolerance)  # warm: smoother calibration, code objects
                t0 = time.perf_counter()
                ores = oe.solve(probes=128, tolerance=args.tolerance)
                ow = time.perf_counter() - t0
                entry = {"rows": oe.dims["rows"], "deletion_units": oe.dims["units"],
                         "physical": oe.dims["physical"], "wall_seconds": round(ow, 4),
                         "probe_solves_per_s": round(3 * 128 / ow, 1), "corrected": ores["corrected"],
                         "numerical_mcse": ores["numerical_mcse"]}
            finally:
                oe.release()
            observation[mode] = entry
            log(f"{mode} deletion KSS ({sub.name}, 128 probes): {ow:.3f}s, corrected {ores['corrected']}")
        # stayers(both): the match-deletion movers plus the eligible stayers
        he = F.Engine(*ocols)
        try:
            st = synth.eligible_stayers(ocols[0], ocols[1], ocols[3], ocols[4], ocols[5], he.mask())
            rec = he.augment_stayers(*st)
            he.configure(args.precision, 64, 32)
            he.solve(probes=16, tolerance=args.tolerance)
            t0 = time.perf_counter()
            hres = he.solve(probes=128, tolerance=args.tolerance)
            hw = time.perf_counter() - t0
            observation["stayers_both"] = {
                "mover_rows": rec["mover_stored_rows"], "stayer_rows": rec["stayer_stored_rows"],
                "stayer_workers": rec["stayer_workers"], "deletion_units": rec["combined_deletion_units"],
                "wall_seconds": round(hw, 4), "probe_solves_per_s": round(3 * 128 / hw, 1),
                "corrected": hres["corrected"], "numerical_mcse": hres["numerical_mcse"]}
        finally:
            he.release()
        log(f"stayer hybrid KSS ({sub.name}, {rec['stayer_workers']} stayers, 128 probes): {hw:.3f}s, "
            f"corrected {hres['corrected']}")
        if args.cpu_solves > 0:
            # the oracle on the host cores beside them: the same 128-probe KSS
            # estimate (fit + 128 leverage + 256 target solves) per route
            from oracle import oracle as O
            for mode, make in (("observation", lambda: O.Problem(*ocols, observation=True)),
                               ("stayers_both", lambda: O.Problem(*ocols).with_stayers(*st))):
                prob = make()
                t0 = time.perf_counter()
                prob.run(probes=128, leverage_width=8, target_width=8, threads=threads,
                         probe_tolerance=args.tolerance)
                cw = time.perf_counter() - t0
                observation[mode]["cpu_oracle_kss_seconds"] = round(cw, 3)
                observation[mode]["cpu_oracle_probe_solves_per_s"] = round(3 * 128 / cw, 2)
                observation[mode]["gpu_vs_cpu_oracle"] = round(cw / observation[mode]["wall_seconds"], 1)
                log(f"oracle {mode} 128-probe KSS on the host cores: {cw:.2f}s")

    line = {
        "metric": "probe solves/s at rel.res 1e-8 (leverage probes, learned+batched CMG-PCG)",
        "value": round(value, 3),
        "unit": "probe solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall / args.steps * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64 Krylov / f%d V-cycle" % args.precision,
        "data": "synthetic",
        "config": {"workload": spec.name + " leverage probe solves to rel.res %g" % args.tolerance,
                   "workers": d["workers"], "firms": d["firms"], "cells": d["cells"],
                   "deletion_units": d["units"], "hybrid_vertices": d["hybrid_vertices"],
                   "cmg_levels": d["levels"], "rhs_per_batch": args.width,
                   "vcycle_precision": args.precision,
                   "smoother": "learned" if smoother and smoother["active"] else "plain",
                   "cycle_levels": smoother["shape"] if smoother else None,
                   "segment_chunk": tuning["chunk"] if tuning else 64,
                   "iterations_per_solve": round(iters, 2)},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "kernel": "k_segments<double, %d, 1, OpDirectionCurvature<%d, %s>> pass 1 "
                               "(finest direction and A p)"
                               % (args.width, args.width, "float" if args.precision == 32 else "double"),
                     "kernel_us": round(tm["kernel_ms"] * 1e3, 2), "launches": tm["kernel_launches"],
                     "algorithmic_bytes": tm["kernel_bytes"]},
        "vcycle_spmv": {"kernel": "finest prolong + smooth pass 1 (FP%d V-cycle)" % args.precision,
                        "achieved": round(ps_gbs, 1), "frac": round(ps_gbs / HBM_PEAK_GBS, 4),
                        "kernel_us": round(ps["ms"] * 1e3, 2), "launches": ps["launches"],
                        "algorithmic_bytes": ps["bytes"],
                        "traffic": measured_traffic(f"w{args.width}_fp{args.precision}",
                                                    "finest_prolong_smooth_traffic.json")},
        "refinement": refinement,
        "cpu_baseline": cpu,
        "autotune": tuning,
        "learned_smoother": smoother,
        "plain_cmg_batched": plain_batched,
        "speedup_vs_plain_cmg_batched": (round(value / plain_batched["solves_per_s"], 3)
                                         if plain_batched else None),
        "plain_fp64_single_rhs": plain,
        "speedup_vs_plain": round(value / plain["solves_per_s"], 2) if plain else None,
        "kss_estimate": kss,
        "controls_estimate": controls,
        "subsample_batch": batch,
        "exact_estimate": exact,
        "observation_estimate": observation,
        "learned_eval_10M": learned_eval,
        "stream_copy_GBs": round(2 * copy_bytes / copy_ms / 1e6, 1),
        "prepare_seconds": round(prep_s, 2),
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
