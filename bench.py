"""Benchmark: GRN cell-steps/s of the batched on-chip path on one MI355X.

Workload (one "step" = one batched evaluation of the whole population):
4096 random GRN models with 3-8 genes and 2-24 links on a 64x64 grid, 2000
explicit-Euler steps with dt = 0.01, FP32 on the device, MSE fitness against
the output of a hidden 6-gene GRN; inputs are an x-gradient, a y-gradient and a
Gaussian blob.  Synthetic, seeded data.  The dataset and the compiled
programs are resident in device memory before the timed region (a staged
batch); a step launches the batch and copies its results back.

Prints one JSON line.  Besides the contract fields it carries:
  host_inclusive_cell_steps_per_s
                the same batch handed over from host memory every step
                (model compilation and the program upload included)
  roofline      achieved FP32 FLOP/s of the field kernel (algorithmic count
                F = 5L + (L - G_h) + 8G + 7D per cell-step, over the kernel
                time measured with HIP events on the library's stream)
                against the 157.3 TFLOP/s FP32 vector peak
  plain         the reference's calling pattern (one model per launch, one
                synchronisation per model) on a sample of the same models
  cpu_baseline  the FP64 CPU oracle on the host cores, bounded sample
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "grn-gpu-performance_amd"))

import numpy as np  # noqa: E402

FP32_PEAK_TFLOPS = 157.3
# HBM bytes per batched launch of the headline workload: FETCH_SIZE x 2.0 +
# WRITE_SIZE x 1.0 (scale factors from the copy-kernel calibration) from the
# two rocprofv3 --pmc runs recorded, with their commands and raw counts, in
# this profile (its "traffic_bytes_per_launch:" line).
TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "field_kernel_config1_64x64.md")


def profiled_traffic():
    """The HBM bytes per launch the committed profile measured, or None."""
    try:
        with open(TRAFFIC_PROFILE) as f:
            for line in f:
                if line.startswith("traffic_bytes_per_launch:"):
                    return float(line.split(":", 1)[1])
    except OSError:
        pass
    return None


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def cpu_baseline(w, gpu_res=None, budget_s=22.0):
    """FP64 oracle on the host cores over a bounded sample of the workload.
    gpu_res (optional): the device's results for every model of the workload,
    compared with the oracle's on the timed sample."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import grn_oracle as ora

    threads = max(1, min(16, len(os.sched_getaffinity(0))))
    p = ora.from_simparams(w.params)
    cells = w.params.dim_x * w.params.dim_y
    # calibrate on one model per thread, then size the timed sample
    probe = w.models[:threads]
    t0 = time.perf_counter()
    r = ora.evaluate_batch(p, w.init, w.target, probe, n_threads=threads)
    dt = time.perf_counter() - t0
    n = int(np.clip(len(probe) * budget_s / max(dt, 1e-3), threads, len(w.models)))
    n = max(threads, n // threads * threads)
    sample = w.models[len(probe):len(probe) + n]
    t0 = time.perf_counter()
    r = ora.evaluate_batch(p, w.init, w.target, sample, n_threads=threads)
    dt = time.perf_counter() - t0
    cs = float(np.sum(r["steps"]) * cells)
    parity = None
    if gpu_res is not None:
        g = gpu_res[len(probe):len(probe) + n]
        rel = np.abs(g["raw_error"] - r["raw_error"]) / np.abs(r["raw_error"])
        order_same = bool(np.array_equal(np.argsort(g["raw_error"], kind="stable"),
                                         np.argsort(r["raw_error"], kind="stable")))
        parity = {"models": int(n), "max_rel_fitness": float(rel.max()), "steps_equal":
                  bool(np.array_equal(g["steps"], r["steps"])), "ranking_identical": order_same}
    # one thread on three more models
    one = w.models[len(probe) + n:len(probe) + n + 3]
    t1 = time.perf_counter()
    r1 = ora.evaluate_batch(p, w.init, w.target, one, n_threads=1)
    dt1 = time.perf_counter() - t1
    cpu_model = None
    try:
        with open("/proc/cpuinfo") as f:
            cpu_model = next((l.split(":", 1)[1].strip() for l in f if l.startswith("model name")), None)
    except OSError:
        pass
    return {"value": cs / dt, "unit": "cell-steps/s", "cores": threads, "kind": "port",
            "sample": f"{len(sample)} of the {len(w.models)} models at full 2000 steps, FP64, "
                      f"{threads} threads, {dt:.1f} s",
            "single_thread_value": float(np.sum(r1["steps"]) * cells) / dt1,
            "single_thread_sample": f"{len(one)} models, {dt1:.1f} s", "cpu_model": cpu_model,
            "parity_vs_gpu": parity}


def config0_full(ctx, seed=0):
    """The reference's own CPU-runnable case in full (64 models, 3 genes,
    32x32, 1000 steps): the FP64 oracle on one thread and on the host share
    of cores, the device on the same models (one batch, and one model per
    launch), and the device's results against the oracle's."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import grn
    import grn_oracle as ora
    import workloads

    w = workloads.config0(seed=seed)
    workloads.make_target(ctx, w)
    ds = ctx.add_dataset(w.params, w.init, w.target)
    p = ora.from_simparams(w.params)
    cells = w.params.dim_x * w.params.dim_y
    threads = max(1, min(16, len(os.sched_getaffinity(0))))
    t0 = time.perf_counter()
    r1 = ora.evaluate_batch(p, w.init, w.target, w.models, n_threads=1)
    t1 = time.perf_counter() - t0
    t0 = time.perf_counter()
    rn = ora.evaluate_batch(p, w.init, w.target, w.models, n_threads=threads)
    tn = time.perf_counter() - t0
    cs = float(np.sum(r1["steps"]) * cells)
    g = ctx.evaluate_batch(w.models, datasets=ds)
    t0 = time.perf_counter()
    g = ctx.evaluate_batch(w.models, datasets=ds)
    tb = time.perf_counter() - t0
    t0 = time.perf_counter()
    for m in w.models:
        ctx.evaluate(m, dataset=ds)
    tp = time.perf_counter() - t0
    rel = np.abs(g["raw_error"] - r1["raw_error"]) / np.abs(r1["raw_error"])
    return {"workload": "config 0: 64 models, 3 genes, 2-6 links, 32x32, 1000 steps, dt 0.01, MSE",
            "cell_steps": cs, "cpu_1thread_s": t1, "cpu_1thread_cell_steps_per_s": cs / t1,
            "cpu_threads": threads, "cpu_host_cpus": os.cpu_count(), "cpu_threads_s": tn,
            "cpu_threads_cell_steps_per_s": cs / tn,
            "gpu_batched_s": tb, "gpu_batched_cell_steps_per_s": cs / tb,
            "gpu_plain_s": tp, "gpu_plain_cell_steps_per_s": cs / tp,
            "parity_vs_oracle": {"max_rel_fitness": float(rel.max()),
                                 "steps_equal": bool(np.array_equal(g["steps"], r1["steps"])),
                                 "ranking_identical": bool(np.array_equal(
                                     np.argsort(g["raw_error"], kind="stable"),
                                     np.argsort(r1["raw_error"], kind="stable"))),
                                 "threads_bitwise_equal": bool(np.array_equal(rn["raw_error"], r1["raw_error"]))}}


def measure_islands(ctx, seed, islands, island_size, warmup, steps, tune=True, tune_sample=256, tune_steps=100):
    """Multi-island evolution generations (16 islands x 1024 models, sides
    32-128, 3000 steps): each generation is the simulation and fitness of
    every island's children in one graph-captured batch of persistent
    launches, then selection and ring migration.  Returns a dict of rates."""
    import grn
    import workloads

    w = workloads.config2(seed=seed, n_islands=islands, island_size=island_size)
    ds = workloads.add_island_datasets(ctx, w)
    search = grn.Search(island_datasets=ds, **w.search)
    tuning = []
    if tune:
        # tile and workgroup shapes per size class, timed on the initial population
        models, dsi, _, _ = search.pending()
        t0 = time.perf_counter()
        reports = ctx.autotune(models, datasets=dsi, sample_models=tune_sample, step_cap=tune_steps)
        log(f"[bench] autotuned {len(reports)} size classes in {time.perf_counter() - t0:.1f} s")
        for r in reports:
            c, d = r["cand"][r["chosen"]], r["cand"][0]
            kind = ["on-chip", "on-chip any width", "halo-tiled"][c["kind"]]
            tuning.append({"grid": f"{r['dim_x']}x{r['dim_y']}", "genes": r["genes"], "models": r["n_models"],
                           "chosen": {"kernel": kind, "cells_per_thread": c["cells_per_thread"],
                                      "block": c["block"], "tile": [c["tile_w"], c["tile_h"]],
                                      "interior": [c["tile_sx"], c["tile_sy"]], "halo": c["halo"]},
                           "candidates": r["n_candidates"],
                           "speedup_vs_default": d["ms"] / c["ms"] if d["ms"] > 0 else None})
    ctx.set_launch_mode(grn.LAUNCH_GRAPH)
    log(f"[bench] {w.name}: warmup x{warmup}")
    for _ in range(warmup):
        search.run_generation(ctx)
    log(f"[bench] timing x{steps} generations")
    cell_steps, alg_flops, kernel_ms, replays, gens = 0.0, 0.0, [], 0, []
    t0 = time.perf_counter()
    for _ in range(steps):
        g = search.run_generation(ctx)
        st = ctx.stats()
        cell_steps += st["cell_steps"]
        alg_flops += st["alg_flops"]
        kernel_ms.append(st["kernel_ms"])
        replays += st["graph_replayed"]
        gens.append(g)
    wall = time.perf_counter() - t0
    ctx.set_launch_mode(grn.LAUNCH_EAGER)
    k = float(np.sum(kernel_ms)) / 1e3
    achieved = alg_flops / k / 1e12
    out = {"workload": "multi-island evolution, config 2: " + w.name, "sides": w.sides,
           "cell_steps_per_s": cell_steps / wall, "kernel_cell_steps_per_s": cell_steps / k,
           "ms_per_generation": wall / steps * 1e3, "kernel_ms_per_generation": k / steps * 1e3,
           "alg_tflops": achieved, "frac_of_fp32_peak": achieved / FP32_PEAK_TFLOPS,
           "launches_per_generation": ctx.stats()["n_launches"], "graph_replays": replays,
           "best_fitness": gens[-1]["min_fit"], "autotune": tuning}
    search.close()
    return out


def measure_config4(ctx, seed, n_islands, island_size, screened, warmup=0, steps=1):
    """Generations of the full-evolution workload at its real grid (100x100,
    5000 steps, four inputs, hidden 10-gene target): n_islands x island_size
    models per generation, plain or screened by the learned surrogate (4
    candidates per slot, the best-predicted one simulated; the surrogate
    trained on (child, slot parent) pairs).  Returns a dict."""
    import grn
    import evolve

    prob = evolve.scaled_config4(seed=seed, n_islands=n_islands, island_size=island_size, side=100, steps=5000)
    evolve.make_target(ctx, prob)
    ds = ctx.add_dataset(prob.params, prob.init, prob.target)
    search = grn.Search(island_datasets=[ds] * n_islands, **prob.search)
    # tile and workgroup shapes per size class, timed on the initial population
    models, dsi, _, _ = search.pending()
    t0 = time.perf_counter()
    reports = ctx.autotune(models, datasets=dsi, sample_models=256, step_cap=100)
    tune_s = time.perf_counter() - t0
    sur = None
    if screened:
        sur = grn.Surrogate(0, n_morphogens=prob.search["n_morphogens"], seed=prob.search["seed"])
        sur.set_pair_weight(4.0)
        search.set_oversample(4)
        search.set_screen_policy(grn.SCREEN_ABSOLUTE)
    ctx.set_launch_mode(grn.LAUNCH_GRAPH)
    # generation 0 (the initial population) is simulated unscreened
    log(f"[bench] config 4: {n_islands} x {island_size} models, {'screened' if screened else 'plain'}, generation 0")
    t0 = time.perf_counter()
    g0 = search.run_generation_screened(ctx, sur, train_steps=20)[0] if screened else search.run_generation(ctx)
    gen0_s = time.perf_counter() - t0
    st0 = ctx.stats()
    for _ in range(warmup):
        search.run_generation_screened(ctx, sur, train_steps=20) if screened else search.run_generation(ctx)
    cell_steps, kernel_ms, alg_flops = 0.0, 0.0, 0.0
    t0 = time.perf_counter()
    for k in range(steps):
        log(f"[bench] config 4: timed generation {k + 1}")
        if screened:
            g, sc = search.run_generation_screened(ctx, sur, train_steps=20)
        else:
            g = search.run_generation(ctx)
        st = ctx.stats()
        cell_steps += st["cell_steps"]
        kernel_ms += st["kernel_ms"]
        alg_flops += st["alg_flops"]
    wall = time.perf_counter() - t0
    ctx.set_launch_mode(grn.LAUNCH_EAGER)
    ctx.clear_tuning()
    out = {"models_per_generation": n_islands * island_size, "screened": screened,
           "autotune_s": tune_s, "autotuned_classes": len(reports), "generation0_s": gen0_s, "generation0_cell_steps_per_s": st0["cell_steps"] / gen0_s,
           "ms_per_generation": wall / steps * 1e3, "cell_steps_per_s": cell_steps / wall,
           "kernel_cell_steps_per_s": cell_steps / (kernel_ms / 1e3),
           "alg_tflops": alg_flops / (kernel_ms / 1e3) / 1e12,
           "frac_of_fp32_peak": alg_flops / (kernel_ms / 1e3) / 1e12 / FP32_PEAK_TFLOPS, "best_fitness": g["min_fit"]}
    if sur is not None:
        out["spearman"] = sc["spearman"]
        sur.close()
    search.close()
    return out


def copy_probe(args):
    """The plain copy kernel over --copy-probe-mib MiB (for calibrating the
    HBM byte counters of a rocprofv3 --pmc run of the same command)."""
    if args.copy_probe_mib:
        import grn

        ms = grn.copy_probe(args.copy_probe_mib << 20)
        log(f"[bench] copy probe: {args.copy_probe_mib} MiB read + written in {ms:.3f} ms "
            f"({2 * (args.copy_probe_mib << 20) / ms / 1e6:.0f} GB/s)")


def measure_classes(ctx, seed, islands, island_size):
    """The config-2 initial population split by (grid side, kernel the
    autotuner chose): each group evaluated alone, as one launch, so its
    kernel time, algorithmic TFLOP/s and cell-steps/s belong to that kernel
    configuration.  Uses the context's current tuning."""
    import grn
    import workloads

    w = workloads.config2(seed=seed, n_islands=islands, island_size=island_size)
    ds = workloads.add_island_datasets(ctx, w)
    search = grn.Search(island_datasets=ds, **w.search)
    models, dsi, _, dts = search.pending()
    reports = ctx.autotune(models, datasets=dsi, sample_models=256, step_cap=100)
    choice = {}
    for r in reports:
        c = r["cand"][r["chosen"]]
        choice[(r["dim_x"], r["genes"])] = (r["dim_x"], ["on-chip", "on-chip any width", "halo-tiled"][c["kind"]],
                                            c["genes_lo"], c["genes_hi"], c["cells_per_thread"], c["block"],
                                            c["tile_w"], c["tile_h"], c["halo"])
    side_of = {d: w.sides[i] for i, d in enumerate(ds)}
    genes = [m.info(w.search["n_morphogens"])[0] for m in models]
    groups = {}
    for i, m in enumerate(models):
        groups.setdefault(choice[(side_of[int(dsi[i])], genes[i])], []).append(i)
    out = []
    for key in sorted(groups):
        idx = groups[key]
        sub = [models[i] for i in idx]
        for _ in range(2):  # the second run is the one reported
            ctx.evaluate_batch(sub, datasets=dsi[idx], dt=dts[idx])
        st = ctx.stats()
        tf = st["alg_flops"] / (st["kernel_ms"] / 1e3) / 1e12
        out.append({"side": key[0], "kernel": key[1], "kernel_genes": [key[2], key[3]],
                    "genes": sorted({genes[i] for i in idx}), "cells_per_thread": key[4], "block": key[5],
                    "tile": [key[6], key[7]], "halo": key[8], "models": len(idx), "launches": st["n_launches"],
                    "kernel_ms": st["kernel_ms"], "alg_tflops": tf, "frac_of_fp32_peak": tf / FP32_PEAK_TFLOPS,
                    "kernel_cell_steps_per_s": st["cell_steps"] / (st["kernel_ms"] / 1e3)})
    search.close()
    return out


def bench_islands(args):
    """The multi-island generation as the bench line's workload (one
    generation is one step)."""
    import grn

    ctx = grn.Context(0)
    copy_probe(args)
    r = measure_islands(ctx, args.seed, args.islands, args.island_size, args.warmup, args.steps,
                        tune=not args.no_tune, tune_sample=args.tune_sample, tune_steps=args.tune_steps)
    out = {
        "metric": "GRN cell-steps/s plain vs batched vs surrogate-screened; time-to-fitness",
        "value": r["cell_steps_per_s"],
        "unit": "cell-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": r["ms_per_generation"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic",
        "config": {"workload": r["workload"], "islands": args.islands,
                   "island_size": args.island_size, "sides": r["sides"], "sim_steps": 3000,
                   "parallelism": "one graph-captured batch of persistent launches per generation"},
        "kernel_cell_steps_per_s": r["kernel_cell_steps_per_s"],
        "roofline": {"bound": "valu", "achieved": r["alg_tflops"], "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": r["frac_of_fp32_peak"], "traffic": None,
                     "kernel_ms": r["kernel_ms_per_generation"]},
        "graph_replays": r["graph_replays"],
        "autotune": r["autotune"],
        "launches_per_generation": r["launches_per_generation"],
        "best_fitness": r["best_fitness"],
        "cpu_baseline": None,
    }
    if args.per_class:
        log("[bench] per-class launches of the initial population")
        out["classes"] = measure_classes(ctx, args.seed, args.islands, args.island_size)
    print(json.dumps(out))
    ctx.close()


def bench_config4(args):
    """Generations of the full-evolution workload at its real size (32
    islands x 2048 models, 100x100, 5000 steps), plain and surrogate-screened
    on the same seed: seconds per generation and cell-steps/s.  Time to
    fitness at this size (hundreds of generations of about a minute) is not
    measured here; --workload evolve measures it at a reduced size."""
    import grn

    ctx = grn.Context(0)
    copy_probe(args)
    plain = measure_config4(ctx, args.seed, args.islands4, args.island_size4, False, steps=args.steps)
    log(f"[bench] plain: {plain}")
    scr = measure_config4(ctx, args.seed, args.islands4, args.island_size4, True, steps=args.steps)
    log(f"[bench] screened: {scr}")
    out = {
        "metric": "GRN cell-steps/s plain vs batched vs surrogate-screened; time-to-fitness",
        "value": plain["cell_steps_per_s"],
        "unit": "cell-steps/s",
        "n_gpus": 1, "steps": args.steps, "warmup": 0, "ms_per_step": plain["ms_per_generation"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "fp32", "data": "synthetic",
        "config": {"workload": f"full evolution, config 4: {args.islands4} islands x {args.island_size4}, "
                               "100x100, 5000 steps, hidden 10-gene target",
                   "parallelism": "one graph-captured batch of persistent launches per generation"},
        "plain": plain, "screened": scr,
        "screened_over_plain_generation_time": plain["ms_per_generation"] / scr["ms_per_generation"],
        "cpu_baseline": None,
    }
    print(json.dumps(out))
    ctx.close()


def coarse_schedule(args):
    """--coarse-generations as evolve takes it: a count, or (side, count) phases."""
    v = str(args.coarse_generations)
    if ":" not in v:
        return int(v)
    return [tuple(int(x) for x in ph.split(":")) for ph in v.split(",")]


def bench_time_to_fitness(args):
    """Plain batched vs surrogate-screened evolution on the same problems and
    seeds: wall time until the simulator-verified best error reaches each
    threshold.  One seed is one problem (its own blob inputs and hidden-GRN
    target) and one search seed; single runs vary a lot, so the value is the
    median over the seeds of plain time / screened time to the plain run's
    final best (0 where the screened run never reached it).  The plain run
    has --generations generations; the screened run the same wall time (it
    ends when it reaches the plain run's final best or when that time is
    up), so a schedule of cheaper generations is not cut short by their
    count."""
    import grn
    import evolve

    seeds = [int(x) for x in args.seeds.split(",")] if args.seeds else [args.seed]
    ctx = grn.Context(0)
    per_seed, ratios, sp, mae, phase = [], [], [], [], {"predict_ms": [], "simulate_ms": [], "record_ms": [],
                                                        "breed_ms": [], "train_ms": []}
    gen_ms = {"plain": [], "screened": []}
    for seed in seeds:
        prob = evolve.scaled_config4(seed=seed)
        evolve.make_target(ctx, prob)
        runs = {}
        for name, screened in (("plain", False), ("screened", args.screen != "none")):
            first = name == "plain"
            log(f"[bench] seed {seed}: {name} evolution, " +
                (f"{args.generations} generations" if first else f"{runs['plain'].wall[-1]:.1f} s"))
            runs[name] = evolve.evolve(ctx, prob, screened,
                                       max_generations=args.generations if first else 20 * args.generations,
                                       threshold=1e-4 if first else runs["plain"].best[-1],
                                       time_limit=None if first else runs["plain"].wall[-1],
                                       oversample=args.oversample, train_steps=args.train_steps,
                                       screen_fraction=args.screen_fraction, screen_policy=args.screen_policy,
                                       screen=args.screen if screened else "mlp", coarse_side=args.coarse_side,
                                       pair_weight=args.pair_weight if screened else 0.0,
                                       coarse_generations=coarse_schedule(args) if name == "screened" else 0,
                                       progress=log)
        plain, scr = runs["plain"], runs["screened"]
        thresholds = sorted({0.02, 0.01, 0.005, plain.best[-1]}, reverse=True)
        table = []
        for thr in thresholds:
            tp, ts = plain.time_to(thr), scr.time_to(thr)
            table.append({"threshold": thr, "plain_s": tp, "screened_s": ts,
                          "speedup": (tp / ts) if tp and ts else None})
        tp, ts = plain.time_to(plain.best[-1]), scr.time_to(plain.best[-1])
        ratios.append((tp / ts) if tp and ts else 0.0)
        screened_gens = [x for x in scr.screen if "coarse_best" not in x]  # not the coarse-to-fine schedule's
        coarse_gens = len(scr.screen) - len(screened_gens)
        if args.screen == "mlp":
            sp += [x["spearman"] for x in screened_gens if x["n_candidates"]]
            mae += [x["mae"] for x in screened_gens if x["n_candidates"]]
            for k in ("predict_ms", "simulate_ms", "record_ms", "breed_ms", "train_ms"):
                phase[k] += [x[k] for x in screened_gens[2:]]
        elif args.screen == "coarse":
            phase.setdefault("coarse_ms", []).extend(x["coarse_ms"] for x in screened_gens[2:])
        if coarse_gens:
            phase.setdefault("coarse_generation_ms", []).append(scr.wall[coarse_gens - 1] / coarse_gens * 1e3)
        gen_ms["plain"].append(plain.wall[-1] / len(plain.wall) * 1e3)
        gen_ms["screened"].append(scr.wall[-1] / len(scr.wall) * 1e3)
        per_seed.append({"seed": seed, "thresholds": table,
                         "plain": {"generations": len(plain.best), "wall_s": plain.wall[-1], "best": plain.best[-1],
                                   "verified": plain.verified},
                         "screened": {"generations": len(scr.best), "wall_s": scr.wall[-1], "best": scr.best[-1],
                                      "verified": scr.verified}})
    out = {
        "metric": "GRN cell-steps/s plain vs batched vs surrogate-screened; time-to-fitness",
        "value": float(np.median(ratios)),
        "unit": "x (median over seeds of plain wall time / screened wall time to the plain run's final "
                "simulator-verified best)",
        "n_gpus": 1,
        "steps": args.generations,
        "warmup": 0,
        "ms_per_step": float(np.mean(gen_ms["plain"])),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32 simulation, bf16/fp32 surrogate",
        "data": "synthetic",
        "config": {"workload": "time to fitness, config 4 at reduced size: " + evolve.scaled_config4().name,
                   "seeds": seeds, "oversample": args.oversample, "screen_fraction": args.screen_fraction,
                   "screen_policy": ["absolute", "delta", "better-half"][args.screen_policy],
                   "screen": args.screen if args.screen != "coarse" else
                   f"coarse simulation on {args.coarse_side}x{args.coarse_side}",
                   "coarse_generations": args.coarse_generations,
                   "train_steps_per_generation": args.train_steps if args.screen == "mlp" else None,
                   "pair_weight": args.pair_weight if args.screen == "mlp" else None},
        "speedup_per_seed": ratios,
        "ms_per_generation": {k: float(np.mean(v)) for k, v in gen_ms.items()},
        "screened_phase_ms": {k: float(np.mean(v)) for k, v in phase.items() if v},
        "spearman_mean": float(np.mean(sp)) if sp else None,
        "mae_mean": float(np.mean(mae)) if mae else None,
        "per_seed": per_seed,
        "cpu_baseline": None,
    }
    print(json.dumps(out))
    ctx.close()


BF16_DENSE_PEAK_TFLOPS = 2500.0


def bench_surrogate(args):
    """The surrogate's BF16-MFMA kernels on their own: Adam training steps
    (batch 4096: gather, 4 forward GEMMs, 7 backward GEMMs, reductions, Adam)
    and inference over 131072 encoded candidates (8 x 16384 offspring)."""
    import grn
    import workloads

    rng = np.random.default_rng(args.seed)
    sur = grn.Surrogate(0, n_morphogens=4, batch=4096, capacity=1 << 16, seed=args.seed)
    models = workloads._population(rng, 4096, (4, 12), (1, 40), 4)
    sur.add(models, rng.uniform(1e-3, 1.0, len(models)))
    log("[bench] surrogate warmup")
    sur.train(max(args.warmup, 1) * 5)
    n_steps = args.steps * 20
    t0 = time.perf_counter()
    sur.train(n_steps)
    train_wall = time.perf_counter() - t0
    st = sur.stats()
    mat = grn.MLP_WIDTH * grn.MLP_WIDTH
    # one step: 4 forward GEMMs + 4 weight-gradient GEMMs + 3 input-gradient GEMMs
    step_flops = 11 * 2.0 * 4096 * mat
    feats = rng.uniform(-1, 1, (131072, grn.MLP_FEATURES)).astype(np.float32)
    sur.predict_features(feats[:8192])
    t0 = time.perf_counter()
    sur.predict_features(feats)
    pred_wall = time.perf_counter() - t0
    pred_ms = sur.stats()["predict_ms"]
    pred_flops = 4 * 2.0 * 131072 * mat
    train_tflops = step_flops * n_steps / (st["train_ms"] * 1e-3) / 1e12
    out = {
        "metric": "GRN cell-steps/s plain vs batched vs surrogate-screened; time-to-fitness",
        "value": 131072 / (pred_ms * 1e-3),
        "unit": "candidates scored/s (device time, 131072 candidates)",
        "n_gpus": 1, "steps": n_steps, "warmup": 5, "ms_per_step": st["train_ms"] / n_steps,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "bf16 (fp32 accumulate)",
        "data": "synthetic",
        "config": {"workload": "surrogate MLP 468 -> 4 x 512 GELU -> 1, Adam batch 4096; inference 131072 rows"},
        "train": {"ms_per_step": st["train_ms"] / n_steps, "wall_ms_per_step": train_wall / n_steps * 1e3,
                  "samples_per_s": 4096 * n_steps / (st["train_ms"] * 1e-3), "tflops": train_tflops},
        "predict": {"device_ms": pred_ms, "wall_ms": pred_wall * 1e3,
                    "tflops": pred_flops / (pred_ms * 1e-3) / 1e12},
        "roofline": {"bound": "mfma", "achieved": train_tflops, "peak": BF16_DENSE_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": train_tflops / BF16_DENSE_PEAK_TFLOPS, "traffic": None},
        "cpu_baseline": None,
    }
    print(json.dumps(out))
    sur.close()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--models", type=int, default=4096)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--seeds", default=",".join(str(i) for i in range(16)),
                    help="evolve: comma-separated problem/search seeds")
    ap.add_argument("--screen-policy", type=int, default=0, help="evolve: 0 absolute, 1 delta, 2 random among the better half")
    ap.add_argument("--screen", choices=["mlp", "coarse", "none"], default="mlp",
                    help="evolve: score candidates with the learned surrogate or a coarse-grid simulation "
                         "(none: the second run is plain, with --coarse-generations)")
    ap.add_argument("--coarse-side", type=int, default=8,
                    help="evolve: side of the coarse grid (--screen coarse, --coarse-generations)")
    ap.add_argument("--coarse-generations", default="8:100,16:100",
                    help="evolve: the second run's first generations simulate on coarse grids only "
                         "(a coarse-to-fine schedule): a count on --coarse-side, or side:count phases; "
                         "0 for none")
    ap.add_argument("--plain-sample", type=int, default=16)
    ap.add_argument("--path", choices=["auto", "lds", "tiled"], default="auto",
                    help="config1: kernel path of the batch (tiled: force the halo-tiled kernels, for comparison)")
    ap.add_argument("--sim-steps", type=int, default=0, help="config1: Euler steps (0: the workload's 2000)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extra", action="store_true",
                    help="config1: skip the config-2 generation and config-4 island reported beside the headline")
    ap.add_argument("--islands4", type=int, default=32, help="config4: islands")
    ap.add_argument("--island-size4", type=int, default=2048, help="config4: models per island")
    ap.add_argument("--workload", choices=["config1", "config2", "config4", "evolve", "surrogate"], default="config1",
                    help="config1: one batch of 4096 models on 64x64 (the headline line); config2: "
                         "multi-island evolution generations on mixed grid sides; evolve: time to "
                         "fitness, plain vs surrogate-screened; surrogate: the MLP's training and "
                         "inference kernels")
    ap.add_argument("--generations", type=int, default=400)
    ap.add_argument("--oversample", type=int, default=4)
    ap.add_argument("--screen-fraction", type=float, default=1.0,
                    help="evolve: fraction of each island's slots simulated per generation")
    ap.add_argument("--train-steps", type=int, default=20)
    ap.add_argument("--pair-weight", type=float, default=4.0,
                    help="evolve: > 0 trains the surrogate on (child, slot parent) pairs as well")
    ap.add_argument("--no-tune", action="store_true", help="config2: skip the autotuner")
    ap.add_argument("--per-class", action="store_true",
                    help="config2: also time each (grid side, chosen kernel) group of the initial population alone")
    ap.add_argument("--tune", action="store_true", help="config1: autotune the size classes first")
    ap.add_argument("--tune-sample", type=int, default=256)
    ap.add_argument("--tune-steps", type=int, default=100)
    ap.add_argument("--islands", type=int, default=16)
    ap.add_argument("--island-size", type=int, default=1024)
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="after the run, copy this many MiB device-to-device with the library's "
                         "copy kernel (known size, to calibrate FETCH_SIZE / WRITE_SIZE counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported "
              f"(models would shard across GPUs by island)", file=sys.stderr)
        sys.exit(2)

    if args.workload == "config2":
        return bench_islands(args)
    if args.workload == "config4":
        return bench_config4(args)
    if args.workload == "evolve":
        return bench_time_to_fitness(args)
    if args.workload == "surrogate":
        return bench_surrogate(args)

    import grn
    import workloads

    w = workloads.config1(seed=args.seed, n_models=args.models)
    if args.sim_steps > 0:
        w.params.num_sim_steps = args.sim_steps
    path = {"auto": grn.PATH_AUTO, "lds": grn.PATH_LDS, "tiled": grn.PATH_TILED}[args.path]
    ctx = grn.Context(0)
    workloads.make_target(ctx, w)
    ds = ctx.add_dataset(w.params, w.init, w.target)
    cells = w.params.dim_x * w.params.dim_y
    flops_per_model = np.array([m.info(4)[2] for m in w.models])

    batch = grn.ModelBatch(w.models)  # the population, marshalled once as a C caller holds it
    tuning = None
    if args.tune:
        reports = ctx.autotune(batch, datasets=ds, sample_models=args.tune_sample, step_cap=args.tune_steps)
        tuning = [{"genes": r["genes"], "models": r["n_models"],
                   "chosen": {k: r["cand"][r["chosen"]][k] for k in ("kind", "cells_per_thread", "block")},
                   "speedup_vs_default": (r["cand"][0]["ms"] / r["cand"][r["chosen"]]["ms"])
                   if r["cand"][0]["ms"] > 0 else None} for r in reports]
        log(f"[bench] autotuned {len(reports)} size classes")
    # the programs are compiled and uploaded before the timed region: a step
    # launches the resident batch and collects its results
    staged = ctx.stage_batch(batch, datasets=ds, path=path)
    log(f"[bench] warmup x{args.warmup}")
    for _ in range(args.warmup):
        staged.run()
    log(f"[bench] timing x{args.steps}")
    kernel_ms, host_ms, cell_steps, alg_flops = [], [], 0.0, 0.0
    t0 = time.perf_counter()
    for _ in range(args.steps):
        res = staged.run()
        st = ctx.stats()
        kernel_ms.append(st["kernel_ms"])
        host_ms.append(st["host_ms"])
        cell_steps += st["cell_steps"]
        alg_flops += st["alg_flops"]
    wall = time.perf_counter() - t0
    assert np.all(res["steps"] == w.params.num_sim_steps)
    value = cell_steps / wall
    # the same batch handed over from host memory every step (model
    # compilation and the PCIe upload of the programs included)
    log(f"[bench] host-inclusive timing x{args.steps}")
    t0 = time.perf_counter()
    for _ in range(args.steps):
        res_h = ctx.evaluate_batch(batch, datasets=ds, path=path)
    host_inclusive = cell_steps / (time.perf_counter() - t0)
    assert np.array_equal(res_h, res)
    staged.close()
    k_avg = float(np.mean(kernel_ms)) / 1e3
    achieved = (alg_flops / args.steps) / k_avg / 1e12

    log(f"[bench] plain path on {args.plain_sample} models")
    sample = w.models[: args.plain_sample]
    plain_rate = None
    if sample:
        t0 = time.perf_counter()
        for m in sample:
            ctx.evaluate(m, dataset=ds)
        plain_wall = time.perf_counter() - t0
        plain_rate = len(sample) * cells * w.params.num_sim_steps / plain_wall

    copy_probe(args)

    cpu = None
    if not args.no_cpu:
        log("[bench] CPU baseline")
        cpu = cpu_baseline(w, gpu_res=res)
        log("[bench] config 0 in full on the CPU oracle and the device")
        cpu["config0_full"] = config0_full(ctx, args.seed)

    # the other single-GPU workloads at their real sizes (reported beside the
    # headline, not in `value`)
    extra = {}
    if not args.no_extra:
        log("[bench] multi-island generation (16 islands x 1024, sides 32-128, 3000 steps)")
        extra["config2_generation"] = measure_islands(ctx, args.seed, 16, 1024, 1, 1)
        log("[bench] one island of the full-evolution workload (2048 models, 100x100, 5000 steps)")
        extra["config4_island"] = measure_config4(ctx, args.seed, 1, 2048, False, steps=1)

    out = {
        "metric": "GRN cell-steps/s plain vs batched vs surrogate-screened; time-to-fitness",
        "value": value,
        "unit": "cell-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic",
        "config": {"workload": "batched GRN simulation+fitness, config 1: 4096 models, 3-8 genes, "
                               "2-24 links, 64x64 grid, 2000 Euler steps, dt 0.01, MSE vs hidden 6-gene GRN",
                   "models": args.models, "grid": "64x64", "sim_steps": w.params.num_sim_steps,
                   "parallelism": "one persistent launch per batch"},
        "kernel_cell_steps_per_s": cell_steps / args.steps / k_avg,
        "host_inclusive_cell_steps_per_s": host_inclusive,
        "roofline": {"bound": "valu", "achieved": achieved, "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / FP32_PEAK_TFLOPS,
                     "traffic": profiled_traffic() if (args.models == 4096 and args.path == "auto"
                                                        and args.sim_steps in (0, 2000)) else None,
                     "kernel_ms": float(np.mean(kernel_ms)),
                     "flops_per_cell_step_mean": float(np.mean(flops_per_model))},
        "autotune": tuning,
        "plain": {"value": plain_rate, "unit": "cell-steps/s", "models": len(sample),
                  "pattern": "one model per launch, synchronise per model"},
        "batched_over_plain": value / plain_rate if plain_rate else None,
        "parity_max_rel": cpu["parity_vs_gpu"]["max_rel_fitness"] if cpu else None,
        "cpu_baseline": cpu,
    }
    out.update(extra)
    print(json.dumps(out))
    ctx.close()


if __name__ == "__main__":
    main()
