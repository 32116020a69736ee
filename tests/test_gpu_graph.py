"""Graph-captured batches (grn_set_launch_mode(GRN_LAUNCH_GRAPH)).

A batch's launches (result clearing plus one persistent kernel per class)
are captured once per layout and replayed; each launch reads its work queue
from a device table uploaded per batch.  Results must equal the eager
launches' bit for bit on every replay, with eager work on the same context
and other batches in between, and a search driven through captured
generations must follow the eager search exactly (same seeds).
"""
import numpy as np
import pytest

import grn
import workloads

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    print("\n[gpu] creating context on device 0", flush=True)
    c = grn.Context(0)
    yield c
    c.set_launch_mode(grn.LAUNCH_EAGER)
    c.close()


def mixed_batch(ctx, seed=0, steps=200):
    """Models on three grids (two on-chip widths and one halo-tiled side)."""
    rng = np.random.default_rng(seed)
    ids, models = [], []
    for s in (32, 48, 96):
        p = grn.SimParams(s, s, 2, 1, steps, 0.05, fitness_kind=grn.FITNESS_MSE)
        init = np.stack([workloads.gradient_x(s, s), workloads.gradient_y(s, s)])[None]
        target = rng.uniform(0, 1, (1, 1, s, s))
        ds = ctx.add_dataset(p, init, target)
        ms = workloads._population(rng, 24, (3, 12), (1, 30), 3)
        ids += [ds] * len(ms)
        models += ms
    return ids, models


def test_graph_replays_equal_eager(ctx):
    ids, models = mixed_batch(ctx)
    ids2, models2 = mixed_batch(ctx, seed=1)
    ctx.set_launch_mode(grn.LAUNCH_EAGER)
    want = ctx.evaluate_batch(models, datasets=ids)
    want2 = ctx.evaluate_batch(models2, datasets=ids2)
    ctx.set_launch_mode(grn.LAUNCH_GRAPH)
    for rep in range(4):
        print(f"[gpu] graph replay {rep}", flush=True)
        got = ctx.evaluate_batch(models, datasets=ids)
        st = ctx.stats()
        assert st["launch_mode"] == grn.LAUNCH_GRAPH and st["n_launches"] >= 3
        assert np.array_equal(got, want)
        if rep > 0:
            assert st["graph_replayed"] == 1
        # eager work and a different batch between replays
        ctx.set_launch_mode(grn.LAUNCH_EAGER)
        assert np.array_equal(ctx.evaluate_batch(models2, datasets=ids2), want2)
        ctx.set_launch_mode(grn.LAUNCH_GRAPH)
        assert np.array_equal(ctx.evaluate_batch(models2, datasets=ids2), want2)
    # a smaller batch (its own layout) keeps every model's result
    sub = [0, 1, 30, 31, 60, 61]
    got = ctx.evaluate_batch([models[i] for i in sub], datasets=[ids[i] for i in sub])
    assert np.array_equal(got, want[sub])
    ctx.set_launch_mode(grn.LAUNCH_EAGER)


def test_search_through_captured_generations_equals_eager(ctx):
    """Two identical searches, one evaluated through captured graphs and one
    eagerly: every generation the per-model results of the two evaluations
    are compared first (a kernel-side difference shows up here), then the
    generation statistics and the bred children (a host-side one)."""
    w = workloads.config2(seed=3, n_islands=3, island_size=48, steps=200, sides=[32, 48, 96])
    ds = workloads.add_island_datasets(ctx, w)
    params = dict(w.search, migration_period=2)
    a = grn.Search(island_datasets=ds, **params)
    b = grn.Search(island_datasets=ds, **params)
    replays = 0
    for gen in range(6):
        print(f"[gpu] search generation {gen}", flush=True)
        ma, da, ea, ta = a.pending()
        mb, db, eb, tb = b.pending()
        assert [m.to_text(17) for m in ma] == [m.to_text(17) for m in mb]
        assert np.array_equal(da, db) and np.array_equal(ea, eb) and np.array_equal(ta, tb)
        ctx.set_launch_mode(grn.LAUNCH_GRAPH)
        ra = ctx.evaluate_batch(ma, datasets=da, max_errors=ea, dt=ta)
        replays += ctx.stats()["graph_replayed"]
        ctx.set_launch_mode(grn.LAUNCH_EAGER)
        rb = ctx.evaluate_batch(mb, datasets=db, max_errors=eb, dt=tb)
        assert np.array_equal(ra["raw_error"], rb["raw_error"]), gen
        assert np.array_equal(ra, rb)
        sa = a.advance(ra["fitness"])
        sb = b.advance(rb["fitness"])
        assert sa == sb
    assert a.best()[1] == b.best()[1]
    print(f"[gpu] graph replays: {replays}", flush=True)


def test_generation_with_host_stage_in_the_graph_equals_eager(ctx):
    """grn_search_run_generation: selection, the Pareto front, statistics,
    ring migration and breeding run as the host stage of the generation's
    batch -- a host node of the captured graph in graph mode -- so a
    generation is one graph launch.  Driven that way, a search follows the
    same search run eagerly, and the search driven through the phased calls
    (pending, evaluate_batch, advance) exactly."""
    # one grid side and at most 8 genes: every generation is one kernel class
    # with the same launch layout, so each one after the first replays the graph
    w = workloads.config2(seed=4, n_islands=3, island_size=96, steps=150, sides=[32, 32, 32])
    ds = workloads.add_island_datasets(ctx, w)
    params = dict(w.search, migration_period=2, n_max_products=8, init_products_max=8)
    a = grn.Search(island_datasets=ds, **params)
    b = grn.Search(island_datasets=ds, **params)
    c = grn.Search(island_datasets=ds, **params)
    replays = 0
    for gen in range(7):
        print(f"[gpu] generation {gen} with its host stage", flush=True)
        ctx.set_launch_mode(grn.LAUNCH_GRAPH)
        sa = a.run_generation(ctx)
        replays += ctx.stats()["graph_replayed"]
        ctx.set_launch_mode(grn.LAUNCH_EAGER)
        sb = b.run_generation(ctx)
        mc, dc, ec, tc = c.pending()
        rc = ctx.evaluate_batch(mc, datasets=dc, max_errors=ec, dt=tc)
        sc = c.advance(rc["fitness"])
        assert sa == sb == sc, gen
        pa, pb, pc = a.pending()[0], b.pending()[0], c.pending()[0]
        assert [m.to_text(17) for m in pa] == [m.to_text(17) for m in pb] == [m.to_text(17) for m in pc]
    assert replays == 6
    assert [x["error"] for x in a.pareto()] == [x["error"] for x in c.pareto()]


def test_context_close_frees_staged_batches():
    """Closing a context with a staged batch still alive frees the batch
    first; a handle freed with its context is ignored by
    grn_staged_destroy and rejected by grn_run_staged."""
    c = grn.Context(0)
    ids, models = mixed_batch(c, seed=7)
    staged = c.stage_batch(models[:4], datasets=ids[:4])
    raw = staged._h
    c.close()
    assert staged._h is None
    lib = grn.load_library()
    lib.grn_staged_destroy(raw)  # already freed: ignored
    # the same through the C ABI alone: grn_destroy with a batch still staged
    c2 = grn.Context(0)
    ids2, models2 = mixed_batch(c2, seed=8)
    st2 = c2.stage_batch(models2[:4], datasets=ids2[:4])
    h2 = st2._h
    st2._h = None  # the wrapper forgets it; only the C side frees it
    c2._staged.discard(st2)
    out = np.zeros(4, grn.RESULT_DTYPE)
    other = grn.Context(0)
    lib.grn_destroy(c2._h)
    c2._h = None
    assert lib.grn_run_staged(other._h, h2, out.ctypes.data) == -1  # GRN_E_INVALID
    lib.grn_staged_destroy(h2)  # freed by grn_destroy: ignored
    other.close()


def test_staged_batch_equals_evaluate_batch(ctx):
    """grn_stage_batch / grn_run_staged: a batch uploaded once and launched
    repeatedly (the clear kernel rewinds every class's work queue) returns
    exactly what grn_evaluate_batch returns, eagerly and as a replayed graph,
    with other batches and a new dataset upload in between."""
    ids, models = mixed_batch(ctx, seed=3)
    ids2, models2 = mixed_batch(ctx, seed=4)
    ctx.set_launch_mode(grn.LAUNCH_EAGER)
    want = ctx.evaluate_batch(models, datasets=ids)
    want2 = ctx.evaluate_batch(models2, datasets=ids2)
    staged = ctx.stage_batch(models, datasets=ids)
    for mode in (grn.LAUNCH_EAGER, grn.LAUNCH_GRAPH):
        for rep in range(3):
            print(f"[gpu] staged run, mode {mode}, repeat {rep}", flush=True)
            ctx.set_launch_mode(mode)
            got = staged.run()
            assert np.array_equal(got, want)
            assert ctx.stats()["n_launches"] >= 3
            ctx.set_launch_mode(grn.LAUNCH_EAGER)
            assert np.array_equal(ctx.evaluate_batch(models2, datasets=ids2), want2)
    mixed_batch(ctx, seed=5)  # more datasets: the device dataset table may move
    ctx.set_launch_mode(grn.LAUNCH_EAGER)
    assert np.array_equal(staged.run(), want)
    staged.close()


def test_staged_batch_argument_errors(ctx):
    """Error paths of the staged batch: an empty batch and a batch staged on
    another context are rejected with the library's error, nothing runs."""
    ids, models = mixed_batch(ctx, seed=6)
    with pytest.raises(grn.GrnError):
        ctx.stage_batch([], datasets=[])
    other = grn.Context(0)
    staged = ctx.stage_batch(models[:4], datasets=ids[:4])
    out = np.zeros(4, grn.RESULT_DTYPE)
    rc = other._lib.grn_run_staged(other._h, staged._h, out.ctypes.data)
    assert rc == -1  # GRN_E_INVALID
    want = ctx.evaluate_batch(models[:4], datasets=ids[:4])
    assert np.array_equal(staged.run(), want)
    staged.close()
    other.close()
