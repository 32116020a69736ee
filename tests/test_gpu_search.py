"""The search driven by the HIP path (grn_search_run_generation).

* run_generation == pending + grn_evaluate_batch + advance, bit for bit, so
  the one-call generation adds nothing to the evaluation it wraps;
* every generation a sample of the pending children (with their per-model
  time steps) is re-evaluated by the FP64 oracle: raw error within 1e-5
  relative;
* deterministic crowding keeps the best error from rising, and the reported
  best model re-evaluates to the reported error.
"""
import numpy as np
import pytest

import grn
import grn_oracle as ora
import workloads

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    print("\n[gpu] creating context on device 0", flush=True)
    c = grn.Context(0)
    yield c
    c.close()


def small_problem(ctx, W, H, seed=0, steps=300):
    rng = np.random.default_rng(seed)
    p = grn.SimParams(W, H, 2, 1, steps, 0.05, fitness_kind=grn.FITNESS_MSE)
    init = np.stack([workloads.gradient_x(W, H), workloads.gradient_y(W, H)])[None]
    target = np.zeros((1, 1, H, W))
    target[0, 0] = workloads.gaussian_blob(W, H, rng.uniform(0, W - 1), rng.uniform(0, H - 1), W / 6)
    ds = ctx.add_dataset(p, init, target)
    return p, init, target, ds


SEARCH = dict(n_morphogens=3, n_max_products=8, n_max_links=16, init_products_min=3,
              init_products_max=5, migration=grn.MIGRATION_RING, migration_period=2, dif_hi=0.1,
              dt_rule=1, dt_scale=0.01, dt_cap=0.25)


def test_run_generation_matches_phased_calls_and_oracle(ctx):
    p, init, target, ds = small_problem(ctx, 32, 32)
    p2, init2, target2, ds2 = small_problem(ctx, 48, 40, seed=1)
    islands = [ds, ds2]
    a = grn.Search(island_datasets=islands, n_islands=2, island_size=64, seed=5, **SEARCH)
    b = grn.Search(island_datasets=islands, n_islands=2, island_size=64, seed=5, **SEARCH)
    probs = {ds: (p, init, target), ds2: (p2, init2, target2)}
    rng = np.random.default_rng(0)
    prev = None
    for gen in range(4):
        print(f"[gpu] search generation {gen}", flush=True)
        models, dsi, me, dt = b.pending()
        res = ctx.evaluate_batch(models, datasets=dsi, max_errors=me, dt=dt)
        sb = b.advance(res["fitness"])
        sa = a.run_generation(ctx)
        assert sa == sb
        # oracle re-evaluation of a sample, with the per-model time step
        for i in rng.choice(len(models), 6, replace=False):
            pp, ii, tt = probs[int(dsi[i])]
            d = dt[i] if dt[i] > 0 else pp.dt
            want = ora.evaluate_batch(ora.from_simparams(pp), ii, tt, [models[i]], dt=[d])[0]
            got = ctx.evaluate_batch([models[i]], datasets=int(dsi[i]), dt=[d])[0]
            assert abs(got["raw_error"] - want["raw_error"]) <= 1e-5 * abs(want["raw_error"]) + 1e-300
        if prev is not None:
            assert sa["min_fit"] <= prev["min_fit"]
        prev = sa
    best, err = a.best()
    # the best individual was measured at its own dt on the dataset of the
    # island it lived on then (ring migration may have moved it since)
    dmax = float(np.max(best.genes["dif"]))
    dtb = min(0.01 / dmax, 0.25) if dmax > 0 else 0.25
    assert any(ctx.evaluate_batch([best], datasets=d, dt=[dtb])["fitness"][0] == err for d in islands)
