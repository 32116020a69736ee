"""The autotuner (grn_autotune / grn_apply_tuning).

* every size class of a mixed batch (on-chip widths and a halo-tiled side)
  gets a report whose chosen candidate is the fastest one that ran;
* tuned evaluations equal the untuned ones (the candidates differ only in
  work decomposition: per-cell arithmetic is the same, the fitness sum's
  order may differ) and agree with the FP64 oracle;
* every candidate of every class, pinned in turn through apply_tuning,
  passes parity with the oracle (raw error within 1e-5 relative).
"""
import numpy as np
import pytest

import grn
import grn_oracle as ora
import workloads

pytestmark = pytest.mark.gpu


def mixed(ctx, seed=0, steps=120):
    rng = np.random.default_rng(seed)
    ids, models, probs = [], [], {}
    for side in (32, 48, 96):
        p = grn.SimParams(side, side, 2, 1, steps, 0.05, fitness_kind=grn.FITNESS_MSE)
        init = np.stack([workloads.gradient_x(side, side), workloads.gradient_y(side, side)])[None]
        target = workloads.gaussian_blob(side, side, side / 3, side / 2, side / 6)[None, None]
        ds = ctx.add_dataset(p, init, target)
        probs[ds] = (p, init, target)
        ms = workloads._population(rng, 40, (3, 12), (2, 24), 3)
        ids += [ds] * len(ms)
        models += ms
    return ids, models, probs


def test_autotune_chooses_fastest_and_every_candidate_passes_parity():
    ctx = grn.Context(0)
    ids, models, probs = mixed(ctx)
    print("\n[gpu] untuned batch", flush=True)
    base = ctx.evaluate_batch(models, datasets=ids)
    want = np.zeros(len(models))
    for ds, (p, init, target) in probs.items():
        sel = [i for i, d in enumerate(ids) if d == ds]
        r = ora.evaluate_batch(ora.from_simparams(p), init, target, [models[i] for i in sel], n_threads=8)
        want[sel] = r["raw_error"]
    assert np.all(np.abs(base["raw_error"] - want) <= 1e-5 * np.abs(want))

    print("[gpu] autotune", flush=True)
    reports = ctx.autotune(models, datasets=ids, sample_models=24, step_cap=40)
    sides = {(r["dim_x"], r["dim_y"]) for r in reports}
    assert sides == {(32, 32), (48, 48), (96, 96)}
    assert sum(r["n_models"] for r in reports) == len(models)
    for r in reports:
        ran = [c for c in r["cand"] if c["ms"] > 0]
        assert ran, r
        assert r["cand"][r["chosen"]]["ms"] == min(c["ms"] for c in ran)
        if r["dim_x"] == 96:
            assert {c["halo"] for c in r["cand"] if c["kind"] == grn.TUNE_TILED} >= {2, 4, 8}

    tuned = ctx.evaluate_batch(models, datasets=ids)
    assert np.array_equal(tuned["steps"], base["steps"])
    assert np.allclose(tuned["raw_error"], base["raw_error"], rtol=1e-12, atol=0)

    print("[gpu] every candidate of every class", flush=True)
    for r in reports:
        cls = [i for i, (d, m) in enumerate(zip(ids, models))
               if probs[d][0].dim_x == r["dim_x"] and m.info(3)[0] == r["genes"]][:4]
        for j in range(r["n_candidates"]):
            if r["cand"][j]["ms"] <= 0:
                continue
            ctx.apply_tuning([dict(r, chosen=j)])
            got = ctx.evaluate_batch([models[i] for i in cls], datasets=[ids[i] for i in cls])["raw_error"]
            assert np.all(np.abs(got - want[cls]) <= 1e-5 * np.abs(want[cls])), (r["dim_x"], r["genes"], j)
    ctx.clear_tuning()
    assert np.allclose(ctx.evaluate_batch(models, datasets=ids)["raw_error"], base["raw_error"], rtol=1e-12)
    ctx.close()


def test_every_tiled_candidate_at_128_passes_parity():
    """128x128 grids take only halo-tiled kernels (with their own window
    geometries: up to 72 cells wide at 8 cells per thread, 1 cell per thread
    at 1024 threads, ...): every candidate of every gene-count class, pinned
    in turn, agrees with the FP64 oracle (raw error within 1e-5 relative)."""
    ctx = grn.Context(0)
    rng = np.random.default_rng(3)
    side, steps = 128, 60
    p = grn.SimParams(side, side, 2, 1, steps, 0.05, fitness_kind=grn.FITNESS_MSE)
    init = np.stack([workloads.gradient_x(side, side), workloads.gradient_y(side, side)])[None]
    target = workloads.gaussian_blob(side, side, side / 3, side / 2, side / 6)[None, None]
    ds = ctx.add_dataset(p, init, target)
    models = workloads._population(rng, 24, (3, 10), (2, 24), 3) + workloads._population(rng, 16, (11, 14), (8, 30), 3)
    models = [m for m in models if m.info(3)[0] <= 12]
    want = ora.evaluate_batch(ora.from_simparams(p), init, target, models, n_threads=8)["raw_error"]
    print("\n[gpu] autotune 128x128", flush=True)
    reports = ctx.autotune(models, datasets=[ds] * len(models), sample_models=8, step_cap=20)
    assert {(r["dim_x"], r["dim_y"]) for r in reports} == {(side, side)}
    kernels = set()
    for r in reports:
        cls = [i for i, m in enumerate(models) if m.info(3)[0] == r["genes"]][:3]
        for j in range(r["n_candidates"]):
            c = r["cand"][j]
            if c["ms"] <= 0:
                continue
            assert c["kind"] == grn.TUNE_TILED
            kernels.add((c["cells_per_thread"], c["genes_lo"], c["genes_hi"]))
            ctx.apply_tuning([dict(r, chosen=j)])
            got = ctx.evaluate_batch([models[i] for i in cls], datasets=ds)["raw_error"]
            assert np.all(np.abs(got - want[cls]) <= 1e-5 * np.abs(want[cls])), (r["genes"], j, c)
    # the larger-window and 1024-thread instantiations of the classes present were among them
    classes = {r["genes"] for r in reports}
    assert {11, 12} & classes, classes
    expected = {k for k in [(8, 2, 4), (6, 2, 5), (5, 6, 8), (3, 11, 12), (4, 11, 12), (2, 2, 5), (2, 6, 8),
                            (2, 9, 10), (1, 9, 12)] if any(k[1] <= g <= k[2] for g in classes)}
    assert expected <= kernels, (expected - kernels, kernels)
    ctx.clear_tuning()
    ctx.close()
