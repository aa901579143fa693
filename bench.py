#!/usr/bin/env python3
"""Benchmark: ms/frame and PCG iterations to relative residual 1e-6 of the implicit solve.

Workload (default): 1024 x 1024 hanging sheet (1,048,576 nodes, 1 mm pitch) with node-face
self-contact, a central 256 x 256 breakable woven region (per-yarn lognormal break strains, seed 7),
top row pinned, two bottom-corner grips pulled apart at 5 mm/step; per frame (one time step
h = 1/120 s): prediction, contact detection, 3 projected-Newton iterations (contact terms, assembly,
Galerkin hierarchy, MG-PCG to rel-res 1e-6 with FP64 residual replacement, FP64 line search), commit,
then the damage field and crossing slip. All state lives in HBM before the timed region; per frame
the host only moves the grip goals (64 bytes).

Three methods run the same frames from the same scene:
  plain    the reference's schedule: FP32 V-cycle with line sweeps on every level (coarsest <= 400
           nodes), the Galerkin hierarchy of the first Newton iteration reused by the later ones;
  mixed    the hierarchy rebuilt at every Newton iteration, the coarsest level (17 x 17, the same
           level count as plain) solved exactly in FP64 (banded Cholesky, dense inverse) instead of
           swept, FP32 operators (FP16 coarse storage needs 4x the iterations with contact), and
           one line colour per family on the levels with lines of at most 513 nodes (every line
           of a family solved at once against the lagged iterate; the fine level keeps 3);
  learned  mixed plus the learned coarse-space fields (BF16 message-passing net on the 65 x 65
           level): an A-optimal warm start from them (or, with --learned-mode 2, deflated flexible
           PCG over them too), kept only when it lowers the residual, plain MG-PCG otherwise; the
           net is trained on the GPU from samples that the warm-up frames' plain solves record (not
           the timed frames).
`value` is the learned method's ms/frame; every method's ms/frame and PCG iterations, and the ratios
against plain, are reported beside it.

Prints one JSON line: the metric, the GPU roofline of the dominant kernel class (HIP-event timed in
an instrumented pass over the same frames) and a CPU baseline (the FP64 C oracle on up to 16 host
threads, and on one, over a bounded sample of the same frame, extrapolated with the GPU's measured
iteration counts).
"""
import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT / "break-that-denim_amd"))

HBM_PEAK_GBS = 8000.0
MIXED_COARSEST_NODES = 400
TRAFFIC_FILE = ROOT / "profiles" / "traffic_sheet1024_solver_kernels.md"


def progress(msg):
    """A progress line on stderr (stdout carries only the result line)."""
    print(f"bench: {msg}", file=sys.stderr, flush=True)


def measured_traffic(kernel_class, size):
    """HBM bytes per launch of a kernel class from the committed rocprofv3 --pmc summary (FETCH_SIZE
    and WRITE_SIZE, scaled by a plain copy's counts), for the workload it was measured on."""
    if size != 1024 or not TRAFFIC_FILE.exists():
        return None
    for line in TRAFFIC_FILE.read_text().splitlines():
        f = line.split()
        if len(f) == 3 and f[0] == "traffic_bytes_per_launch" and f[1] == kernel_class:
            return float(f[2])
    return None


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=4)
    ap.add_argument("--size", type=int, default=1024)
    ap.add_argument("--profile-frames", type=int, default=2)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-autotune", action="store_true", help="keep the default smoother kernel per level")
    ap.add_argument("--pin", action="append", default=[], metavar="LEVEL:U:V",
                    help="after the autotuner, run level LEVEL's u- and v-line passes with kernel variants U and V "
                         "(0 direct, 4 / 8 staged, 16 split); e.g. under a kernel tracer, whose per-dispatch cost "
                         "counts against the two-kernel split variant in the autotuner's timing")
    ap.add_argument("--method", choices=("all", "plain", "mixed", "learned"), default="all")
    ap.add_argument("--epochs", type=int, default=200, help="training epochs of the learned method's net")
    ap.add_argument("--learned-mode", type=int, default=1, choices=(1, 2),
                    help="1: the learned warm start; 2: the warm start and the deflated correction")
    ap.add_argument("--no-contact", action="store_true", help="leave node-face self-contact off")
    ap.add_argument("--workload", choices=("sheet1024", "sheet64", "patches48"), default="sheet1024",
                    help="sheet1024: the headline workload (default); sheet64: the 64 x 64 sheet, eager against "
                         "graph-captured substeps; patches48: 48 uneven patches, one eager substep per patch "
                         "against one graph launch for the batch")
    ap.add_argument("--one-colour-max-line", type=int, default=513, metavar="NODES",
                    help="mixed and learned methods: levels whose lines have at most NODES nodes smooth with one "
                         "line colour per family (bd_set_line_colors); 0 keeps the reference's 3")
    ap.add_argument("--lag-coarsest", action="store_true",
                    help="mixed and learned methods: factor the exact coarsest level once per frame, at its first "
                         "Newton iteration (BD_COARSE_EXACT_LAGGED), instead of at every one (BD_COARSE_EXACT)")
    ap.add_argument("--fine-colours", type=int, default=3, choices=(2, 3),
                    help="mixed and learned methods: line colours per family of the fine level (2: lines two apart "
                         "together, their couplings majorised)")
    ap.add_argument("--batch-threads", type=int, default=16,
                    help="host threads of the patches' concurrent batch (grid+threads)")
    ap.add_argument("--sweeps", type=int, default=1,
                    help="pre- and post-smoothing sweeps per level of the mixed and learned methods (plain: 1)")
    ap.add_argument("--mixed-coarsest", type=int, default=MIXED_COARSEST_NODES, metavar="NODES",
                    help="the mixed and learned methods stop the V-cycle at the first level of at most NODES nodes "
                         "and solve it exactly")
    ap.add_argument("--mixed-precision", choices=("fp16", "fp32"), default="fp32",
                    help="operator storage of the mixed method's coarse levels (fine level and plain method: fp32)")
    ap.add_argument("--smoother-probe", type=int, default=0, metavar="SWEEPS",
                    help="only run SWEEPS forward sweeps of every smoothed level of the mixed method's hierarchy "
                         "(after one frame, the autotuner and any --pin), outside PCG: the smoother passes for PMC "
                         "counters")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only run 8 plain device copies of MIB MiB (calibrates the PMC byte counters)")
    return ap.parse_args()


def copy_probe(mib):
    """A plain device-to-device copy of known size, timed and counted under rocprofv3 --pmc to scale
    FETCH_SIZE / WRITE_SIZE to bytes."""
    import torch
    n = mib * (1 << 20) // 4
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    for _ in range(8):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_bytes_read": 4 * n, "copy_probe_bytes_written": 4 * n, "copies": 8}), flush=True)


def smoother_probe(args, sweeps):
    """The mixed method's smoother passes on their own: one frame of the 1024^2 workload builds the hierarchy, the
    autotuner picks the kernels, then `sweeps` forward sweeps of each smoothed level run through bd_smooth (every
    pass does its full work; inside PCG a converged solve gates the remaining V-cycle kernels)."""
    import numpy as np
    from bd.capi import BD_COARSE_EXACT
    from bd.configs import config_sheet
    from bd.solver import GpuSolver
    cfg = config_sheet(args.size)
    params = cfg["params"]
    params.coarse_solve = BD_COARSE_EXACT
    params.max_coarsest_nodes = MIXED_COARSEST_NODES
    gpu = GpuSolver(cfg["scene"], params, cfg["grips"])
    goals = np.array([g["goal"] for g in cfg["grips"]], dtype=np.float64) + np.array(cfg["velocity"])
    gpu.set_grip_goals(goals)
    gpu.encode_substep()
    tuned = dict(gpu.autotune(3))
    for pin in args.pin:
        lvl, u, v = (int(q) for q in pin.split(":"))
        gpu.set_smoother(lvl, 0, u)
        gpu.set_smoother(lvl, 1, v)
        tuned[lvl] = (u, v)
    levels = gpu.stats()["level_nodes"]
    for k in range(len(levels) - 1):
        for _ in range(sweeps):
            gpu.encode_smooth(k, False)
    gpu.sync()
    print(json.dumps({"smoother_probe_levels": levels[:-1], "sweeps": sweeps,
                      "smoother_kernels": {str(k): list(v) for k, v in tuned.items()}}), flush=True)
    gpu.close()


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(cfg, gpu_iters_per_newton, newton_per_frame):
    """Time the FP64 oracle on one Newton iteration of the same frame (assemble + rhs + Galerkin
    hierarchy + 2 PCG iterations + line search), on one thread and on the threads OpenMP is given
    (the matvec and line-smoother loops are threaded; the results are the same); extrapolate both to
    the GPU's iteration counts. `value` is the threaded one."""
    sys.path.insert(0, str(ROOT / "oracle"))
    import oracle as O
    from bd.configs import advance_grips

    def sample(threads):
        used = O.set_threads(threads)
        grips = [dict(g) for g in cfg["grips"]]
        advance_grips(grips, cfg["velocity"])
        o = O.Oracle(cfg["scene"], cfg["params"], grips)
        o.predict()
        t0 = time.perf_counter()
        o.assemble()
        b = o.rhs()
        o.build_hierarchy()
        t1 = time.perf_counter()
        d = o.pcg(b, 2)
        t2 = time.perf_counter()
        o.energy_change(d)
        t3 = time.perf_counter()
        setup = (t1 - t0) + (t3 - t2)
        per_iter = (t2 - t1) / 2.0
        return used, 1e3 * newton_per_frame * (setup + gpu_iters_per_newton * per_iter), (t1 - t0, t2 - t1, t3 - t2)

    most = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or min(16, os.cpu_count() or 1)
    _, one_ms, one = sample(1)
    cores, many_ms, many = sample(most)
    return {"value": round(many_ms, 1), "unit": "ms/frame", "cores": cores, "kind": "port",
            "single_core_value": round(one_ms, 1), "cpu": cpu_model(),
            "sample": f"FP64 C oracle on one Newton iteration of the same {cfg['scene'].node_count}-node frame "
                      f"(assemble+rhs+hierarchy {many[0]:.2f} s, 2 PCG iterations {many[1]:.2f} s, line search "
                      f"{many[2]:.2f} s on {cores} threads; {one[0]:.2f} / {one[1]:.2f} / {one[2]:.2f} s on one), "
                      f"extrapolated to {newton_per_frame} Newton x {gpu_iters_per_newton:.1f} PCG iterations per frame"}


def method_params(cfg, method, args):
    import copy
    from bd.capi import BD_COARSE_EXACT, BD_COARSE_EXACT_LAGGED, BD_COARSE_SWEEPS, BD_PRECISION_FP16, BD_PRECISION_FP32
    params = copy.copy(cfg["params"])
    # the reference's schedule smooths once per level; the mixed and learned methods take --sweeps
    params.smooth_sweeps = 1 if method == "plain" else args.sweeps
    if method == "plain":
        params.coarse_solve = BD_COARSE_SWEEPS
        params.max_coarsest_nodes = 400
        params.lag_coarse = 1
        params.precision = BD_PRECISION_FP32
    else:
        params.coarse_solve = BD_COARSE_EXACT_LAGGED if args.lag_coarsest else BD_COARSE_EXACT
        params.max_coarsest_nodes = args.mixed_coarsest
        params.lag_coarse = 0
        params.precision = BD_PRECISION_FP16 if args.mixed_precision == "fp16" else BD_PRECISION_FP32
    return params


def run_method(args, cfg, method, profile_frames):
    """Warm-up, then exactly `args.steps` timed frames of one method, then an instrumented pass of
    `profile_frames` frames (per-kernel-class HIP event times). The learned method records training
    samples during its warm-up (plain solves, the net off), trains, and then switches the net on."""
    import numpy as np
    from bd.solver import GpuSolver
    params = method_params(cfg, method, args)
    gpu = GpuSolver(cfg["scene"], params, cfg["grips"])
    train = None
    if method == "learned":
        gpu.learn_record(3 * args.warmup)
    goals = np.array([g["goal"] for g in cfg["grips"]], dtype=np.float64)
    vel = np.array(cfg["velocity"], dtype=np.float64)

    def frame():
        nonlocal goals
        goals = goals + vel
        gpu.set_grip_goals(goals)
        gpu.encode_substep()

    colours = None
    for w in range(args.warmup):
        frame()
        if w == 0:
            colours = [3] * len(gpu.stats()["level_nodes"])
            if method != "plain" and args.one_colour_max_line > 0:
                # the levels whose lines are short enough smooth each family's lines at once
                for k, nodes in enumerate(gpu.stats()["level_nodes"]):
                    if round(nodes ** 0.5) <= args.one_colour_max_line:
                        gpu.set_line_colors(k, 1)
                        colours[k] = 1
            if method != "plain" and args.fine_colours != 3 and colours[0] == 3:
                gpu.set_line_colors(0, args.fine_colours)
                colours[0] = args.fine_colours
    gpu.sync()
    if method == "learned":
        t0 = time.perf_counter()
        losses = gpu.learn_train(args.epochs, lr=3e-3, batch=3, targets=3)
        info = gpu.learn_info()
        gpu.set_learned(args.learned_mode)
        train = {"samples": info["samples"], "net_level_nodes": info["nodes"], "epochs": args.epochs,
                 "loss_first_epoch": round(float(losses[0]), 5), "loss_last_epoch": round(float(losses[-1]), 5),
                 "seconds": round(time.perf_counter() - t0, 2)}
    # the fastest smoother kernel per level and family for this hierarchy (bd_autotune)
    # 15 V-cycles per candidate: the variants of a level differ by a few microseconds per V-cycle, so
    # the fastest of fewer runs picks at random between them
    tuned = gpu.autotune(15) if not args.no_autotune else {}
    for pin in args.pin:
        lvl, u, v = (int(q) for q in pin.split(":"))
        gpu.set_smoother(lvl, 0, u)
        gpu.set_smoother(lvl, 1, v)
        tuned = dict(tuned)
        tuned[lvl] = (u, v)
    iters, newton, relres = [], [], []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        frame()
        st = gpu.stats()
        iters.append(sum(st["pcg_iterations"]))
        newton.append(st["newton_iterations"])
        relres.append(max(st["rel_residual"][:st["newton_iterations"]]))
    gpu.sync()
    t1 = time.perf_counter()
    st = gpu.stats()
    out = {"ms": 1e3 * (t1 - t0) / args.steps, "iters": float(np.mean(iters)), "newton": float(np.mean(newton)),
           "iters_list": [int(i) for i in iters],
           "max_rel_residual": float(max(relres)), "train": train, "learned_fallbacks": st["learned_fallbacks"],
           "levels": gpu.stats()["level_nodes"], "tuned": tuned, "precision": params.precision, "colours": colours,
           "sweeps": params.smooth_sweeps}
    gpu.profile_enable(True)
    for _ in range(profile_frames):
        frame()
    out["prof"] = gpu.profile_read()
    gpu.profile_enable(False)
    out["finite"] = bool(np.isfinite(gpu.read(0)).all())
    gpu.close()
    return out


def graph_budget(iters_per_newton, limit):
    """The PCG budget of a captured Newton solve: 1.5x the most iterations an eager warm-up solve
    took, plus 4, within the solver's own limit."""
    return int(min(limit, max(8, int(1.5 * iters_per_newton) + 4)))


def graph_bench(args):
    """Launch-bound workloads, several ways on the same frames: cfg 1's 64 x 64 sheet (one solver per
    method) or 48 uneven patches (each its own Newton / MG-PCG system). Methods:
      grid            PCG kernels over the chip, the host checking convergence each iteration; the
                      patches one substep after another;
      grid+threads    (the patches) the grid path's eager substeps on each patch's own stream, driven by
                      a pool of host threads (bd_batch_create_concurrent), so that the patches'
                      dependent launches overlap on the chip;
      grid+graph+threads (the patches) the captured substeps replayed on each patch's own stream from the
                      concurrent batch's host threads;
      grid+graph      the substep captured in a graph (the PCG budget per Newton solve from the warm-up:
                      iterations past convergence are launched and return on the device's flag); the
                      patches as the branches of one graph launch;
      workgroup       the whole PCG in one workgroup (BD_PCG_WORKGROUP); the patches in lockstep with
                      each Newton iteration's 48 solves in one launch (fused batch);
      workgroup+graph (the sheet) the one-workgroup substep captured in a graph."""
    import numpy as np
    from bd.capi import BD_PCG_WORKGROUP
    from bd.configs import advance_grips, config1, config_patches
    from bd.solver import Batch, GpuSolver
    single = args.workload == "sheet64"
    if single:
        cfgs = [config1()]
        name = "sheet64 (cfg 1: 64x64 sheet, 3 Newton, 3-level MG-PCG to rel-res 1e-5, one corner grip at 2 mm/step)"
        methods = ("grid", "grid+graph", "workgroup", "workgroup+graph")
    else:
        cfgs = config_patches()["patches"]
        name = (f"patches48 ({len(cfgs)} uneven patches, sides {min(p['side'] for p in cfgs)}-{max(p['side'] for p in cfgs)},"
                f" {sum(p['side'] ** 2 for p in cfgs)} nodes, each 3 Newton x MG-PCG to rel-res 1e-6, random grip paths)")
        methods = ("grid", "grid+threads", "grid+graph", "grid+graph+threads", "workgroup")

    def frame(group, step):
        for g, grips, vel in group:
            advance_grips(grips, vel)
            g.set_grip_goals(np.array([q["goal"] for q in grips]))
            if step:
                step(g)

    def collect(group):
        its, rel = 0, 0.0
        for g, _, _ in group:
            st = g.stats()
            n = st["newton_iterations"]
            its += sum(st["pcg_iterations"][:n])
            rel = max(rel, max(st["rel_residual"][:n]))
        return its, rel

    results, budgets = {}, None
    for m in methods:
        progress(f"method {m}: start")
        group = []
        for c in cfgs:
            grips = [dict(g) for g in c["grips"]]
            g = GpuSolver(c["scene"], c["params"], grips)
            if m.startswith("workgroup"):
                g.set_pcg_mode(BD_PCG_WORKGROUP)
            group.append((g, grips, c["velocity"]))
        most = [0] * len(group)
        for _ in range(args.warmup):
            frame(group, lambda g: g.encode_substep())
            for i, (g, _, _) in enumerate(group):
                st = g.stats()
                most[i] = max(most[i], max(st["pcg_iterations"][:st["newton_iterations"]]))
        batch, step = None, (lambda g: g.encode_substep())
        if "graph" in m:
            caps = [graph_budget(k, c["params"].pcg_max_iterations) for k, c in zip(most, cfgs)]
            if m.startswith("grid+graph"):
                budgets = caps
            for (g, _, _), cap in zip(group, caps):
                g.graph_capture(cap)
            step = (lambda g: g.graph_substep()) if single else None
        if not single and m != "grid":
            batch = Batch([g for g, _, _ in group], threads=args.batch_threads if m.endswith("threads") else None)
            step = None
        for g, _, _ in group:
            g.sync()
        its, rel = 0, 0.0
        t0 = time.perf_counter()
        for _ in range(args.steps):
            frame(group, step)
            if batch:
                batch.substep()
            i, r = collect(group)
            its += i
            rel = max(rel, r)
        for g, _, _ in group:
            g.sync()
        results[m] = {"ms_per_frame": round(1e3 * (time.perf_counter() - t0) / args.steps, 3),
                      "pcg_iterations_per_frame": round(its / args.steps, 1), "max_rel_residual": rel}
        if batch:
            batch.close()
        for g, _, _ in group:
            g.close()
    tol = cfgs[0]["params"].pcg_rel_tol
    main_m = "grid+threads" if "grid+threads" in results else "workgroup"
    ms = results[main_m]["ms_per_frame"]
    print(json.dumps({
        "metric": "ms/frame of launch-bound solves: batched forms against the grid path one system at a time",
        "value": ms, "unit": "ms/frame", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms,
        "higher_is_better": False, "scaling": "weak", "vs_baseline": None,
        "dtype": "fp32 (FP64 dots, residual replacement, line search)", "data": "synthetic",
        "config": {"workload": name, "method": main_m, "graph_pcg_budget_per_newton": budgets,
                   "batch_threads": args.batch_threads if main_m == "grid+threads" else None},
        "methods": results,
        "speedup_grid_over_" + main_m: round(results["grid"]["ms_per_frame"] / ms, 3),
        "converged": {k: v["max_rel_residual"] <= 1.05 * tol for k, v in results.items()},
        "roofline": None, "cpu_baseline": None}), flush=True)


def ratio(runs, a, b, key):
    return round(runs[a][key] / runs[b][key], 3) if a in runs and b in runs and runs[b][key] > 0 else None


def main():
    args = parse()
    if args.copy_probe:
        copy_probe(args.copy_probe)
        return
    if args.smoother_probe:
        smoother_probe(args, args.smoother_probe)
        return
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); the multi-GPU strip decomposition is "
              f"a design note, not part of this build", file=sys.stderr)
        sys.exit(2)
    if args.workload != "sheet1024":
        graph_bench(args)
        return
    from bd.configs import config_sheet

    cfg = config_sheet(args.size, self_contact=not args.no_contact)
    S = cfg["scene"]
    runs = {}
    order = ("plain", "mixed", "learned") if args.method == "all" else (args.method,)
    main_name = order[-1]
    for m in order:
        progress(f"method {m}: start")
        runs[m] = run_method(args, cfg, m, args.profile_frames if m == main_name else 0)
        progress(f"method {m}: {runs[m]['ms']:.1f} ms/frame, {runs[m]['iters']:.1f} PCG iterations per frame")
    R = runs[main_name]
    ms = R["ms"]
    prof = R["prof"]
    hbm = {k: v for k, v in prof.items() if v[2] > 0}
    if not hbm:
        print(json.dumps({"note": "no profiled frames", "ms_per_frame": {k: v["ms"] for k, v in runs.items()}}), flush=True)
        return
    dom = max(hbm, key=lambda k: hbm[k][1])
    launches, tot_ms, nbytes = hbm[dom]
    achieved = nbytes / (tot_ms * 1e-3) / 1e9
    roof = {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": measured_traffic(dom, args.size),
            "avg_launch_us": round(1e3 * tot_ms / launches, 2), "bytes_per_launch": round(nbytes / launches)}
    kernels = {k: {"launches": v[0], "ms_per_frame": round(v[1] / args.profile_frames, 3),
                   "GBps": round(v[2] / (v[1] * 1e-3) / 1e9, 1) if v[1] > 0 and v[2] > 0 else None}
               for k, v in sorted(prof.items(), key=lambda kv: -kv[1][1])}
    iters_per_frame = R["iters"]
    newton_per_frame = R["newton"]
    cpu = None
    if not args.no_cpu_baseline and os.environ.get("BD_BENCH_NO_CPU") is None:
        P = runs.get("plain", R)
        progress("cpu baseline: the FP64 oracle on one Newton iteration, threaded and on one core")
        cpu = cpu_baseline(cfg, P["iters"] / max(P["newton"], 1), P["newton"])
    out = {
        "metric": "ms/frame & PCG iters to rel-res 1e-6: learned+mixed vs plain MG-PCG",
        "value": round(ms, 3),
        "unit": "ms/frame",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 3),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32 (FP64 dots, residual replacement, line search" +
                 (", coarsest solve; fp16 coarse-level operators in the V-cycle" if main_name != "plain" and R["precision"]
                  else ", coarsest solve" if main_name != "plain" else "") +
                 ("; bf16 net on the matrix cores, fp32 accumulation)" if main_name == "learned" else ")"),
        "data": "synthetic",
        "config": {"workload": f"sheet{args.size}x{args.size} FP32 MG-PCG to rel-res 1e-6, "
                               f"central {args.size // 4}x{args.size // 4} breakable yarn region, 2 grips at 5 mm/step"
                               + (", self-contact" if not args.no_contact else ""),
                   "nodes": S.node_count, "levels": R["levels"], "newton_per_frame": newton_per_frame,
                   "method": {"plain": "plain (reference V-cycle schedule)", "mixed": "mixed (FP64 exact coarsest solve)",
                              "learned": "learned+mixed (learned coarse-space correction, FP64 exact coarsest solve)"}[main_name]},
        "pcg_iterations_per_frame": iters_per_frame,
        "pcg_iterations_per_newton": round(iters_per_frame / max(newton_per_frame, 1), 2),
        "methods": {k: {"ms_per_frame": round(v["ms"], 3), "pcg_iterations_per_frame": v["iters"],
                        "smooth_sweeps": v["sweeps"], "vcycle_operators": "fp16" if v["precision"] else "fp32",
                        "max_rel_residual": v["max_rel_residual"], "learned_fallbacks": v["learned_fallbacks"],
                        "training": v["train"], "pcg_iterations_each_frame": v["iters_list"],
                        "line_colours_per_level": v["colours"],
                        "smoother_kernels": {str(lv): list(c) for lv, c in v["tuned"].items()}}
                    for k, v in runs.items()},
        "iteration_ratio_plain_over_learned": ratio(runs, "plain", "learned", "iters"),
        "speedup_plain_over_learned": ratio(runs, "plain", "learned", "ms"),
        "iteration_ratio_mixed_over_learned": ratio(runs, "mixed", "learned", "iters"),
        "speedup_mixed_over_learned": ratio(runs, "mixed", "learned", "ms"),
        "iteration_ratio_plain_over_mixed": ratio(runs, "plain", "mixed", "iters"),
        "speedup_plain_over_mixed": ratio(runs, "plain", "mixed", "ms"),
        "finite": all(r["finite"] for r in runs.values()),
        # every timed Newton solve reached the tolerance (a solve stopped by the iteration cap would not)
        "max_rel_residual": max(r["max_rel_residual"] for r in runs.values()),
        "converged": all(r["max_rel_residual"] <= 1.05e-6 for r in runs.values()),
        "roofline": roof,
        "kernels": kernels,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    if not out["converged"]:
        print(f"bench.py: a timed solve stopped above rel-res 1.05e-6 ({out['max_rel_residual']:.3e})", file=sys.stderr)
        sys.exit(3)


if __name__ == "__main__":
    main()
