"""Oracle parity on the multi-island and the full-evolution workloads at their
real sizes: a seeded sample of each one's initial population, simulated at
the full step count on the device (the kernels the batch planner picks:
on-chip for 32-64, halo-tiled for 96-128 and 100x100) and by the FP64 CPU
oracle with the same per-model time steps.

Bars: exact step counts; fitness within 1e-5 relative and identical ranking
for every model whose fitness is well conditioned at FP32 precision, as the
oracle measures it (parity_util: the fitness moves by at most 1e-6 when
every parameter is perturbed by 1e-7); an ill-conditioned model (a
near-chaotic trajectory at these large time steps) must stay within 20x the
oracle's own perturbation spread.  At least 80 % of the sample must be well
conditioned.  Final fields within 1e-4 relative + 1e-7 per cell for a
well-conditioned model.
"""
import numpy as np
import pytest

import evolve
import grn
import grn_oracle as ora
import workloads
from parity_util import check_conditioned, oracle_with_conditioning

pytestmark = pytest.mark.gpu

FIT_RTOL = 1e-5
FIELD_RTOL, FIELD_ATOL = 1e-4, 1e-7
ORACLE_THREADS = 16


@pytest.fixture(scope="module")
def ctx():
    print("\n[gpu] creating context on device 0", flush=True)
    c = grn.Context(0)
    yield c
    c.close()


def check_fields(got, want):
    assert got.shape == want.shape
    bad = np.abs(got - want) > FIELD_RTOL * np.abs(want) + FIELD_ATOL
    assert not bad.any(), f"{bad.sum()} cells out of tolerance, worst abs {np.abs(got - want).max():.3e}"


def test_config2_initial_population_sample(ctx):
    """16 islands x 1024, sides {32, 48, 64, 96, 128}, dt = 0.2 dx^2 / max D
    (cap 1), two target genes (stripe and spot of the hidden 8-gene GRN),
    3000 steps: three models of every side."""
    w = workloads.config2(seed=0)
    print(f"[gpu] config2 targets for sides {sorted(set(w.sides))}", flush=True)
    ds = workloads.add_island_datasets(ctx, w)
    search = grn.Search(island_datasets=ds, **w.search)
    models, dsi, _, dts = search.pending()
    rng = np.random.default_rng(5)
    per_side = {}
    for island, side in enumerate(w.sides):
        per_side.setdefault(side, island)
    assert sorted(per_side) == [32, 48, 64, 96, 128]
    n_well = n_all = 0
    for side, island in sorted(per_side.items()):
        idx = island * w.search["island_size"] + rng.choice(w.search["island_size"], 4, replace=False)
        sel = [models[i] for i in idx]
        print(f"[gpu] config2 side {side}: models {list(idx)}, genes {[len(m.genes) for m in sel]}", flush=True)
        got = ctx.evaluate_batch(sel, datasets=dsi[idx], dt=dts[idx])
        p = w.params[island]
        want, spread = oracle_with_conditioning(ora.from_simparams(p), w.init[island], w.target[island], sel,
                                                dt=dts[idx], n_threads=ORACLE_THREADS)
        assert np.all(got["steps"] == 3000)
        well, rel = check_conditioned(got, want, spread)
        n_well += int(well.sum())
        n_all += len(sel)
        print(f"[gpu] config2 side {side}: rel fitness err {np.array2string(rel, precision=2)}, "
              f"oracle spread {np.array2string(spread, precision=2)}", flush=True)
        if side in (32, 128):
            k = int(np.argmax(well))  # a well-conditioned model of the side
            assert well[k]
            f = ctx.simulate_fields(sel[k], dataset=int(dsi[idx[k]]), dt=float(dts[idx[k]]))
            _, fo = ora.evaluate(ora.params(p.dim_x, p.dim_y, 2, 2, 3000, float(dts[idx[k]])), w.init[island],
                                 w.target[island], sel[k].genes, sel[k].links, want_fields=True)
            check_fields(f, fo)
    assert n_well >= 0.8 * n_all, (n_well, n_all)
    search.close()


def test_config4_initial_population_sample(ctx):
    """32 islands x 2048, 100x100, four inputs (two blobs, two gradients),
    target of the hidden 10-gene GRN, 5000 steps, at most 10 genes: eight
    models of the initial population (the halo-tiled path)."""
    prob = evolve.scaled_config4(seed=1, n_islands=32, island_size=2048, side=100, steps=5000)
    evolve.make_target(ctx, prob)
    ds = ctx.add_dataset(prob.params, prob.init, prob.target)
    search = grn.Search(island_datasets=[ds] * 32, **prob.search)
    models, dsi, _, dts = search.pending()
    assert len(models) == 32 * 2048
    idx = np.random.default_rng(7).choice(len(models), 8, replace=False)
    sel = [models[i] for i in idx]
    print(f"[gpu] config4 models {list(idx)}, genes {[len(m.genes) for m in sel]}", flush=True)
    got = ctx.evaluate_batch(sel, datasets=dsi[idx], dt=dts[idx])
    want, spread = oracle_with_conditioning(ora.from_simparams(prob.params), prob.init, prob.target, sel,
                                            dt=dts[idx], n_threads=ORACLE_THREADS)
    assert np.all(got["steps"] == 5000)
    well, rel = check_conditioned(got, want, spread)
    print(f"[gpu] config4: rel fitness err {np.array2string(rel, precision=2)}, "
          f"oracle spread {np.array2string(spread, precision=2)}", flush=True)
    assert well.sum() >= 0.8 * len(sel)
    sizes = np.where(well, [len(x.genes) for x in sel], -1)
    k = int(np.argmax(sizes))  # the largest well-conditioned sampled network
    m, j = sel[k], int(idx[k])
    f = ctx.simulate_fields(m, dataset=ds, dt=float(dts[j]))
    p = prob.params
    _, fo = ora.evaluate(ora.params(p.dim_x, p.dim_y, 4, 1, 5000, float(dts[j])), prob.init, prob.target, m.genes,
                         m.links, want_fields=True)
    check_fields(f, fo)
    search.close()
