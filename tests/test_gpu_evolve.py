"""Evolution to a fitness target through the batched path, plain and
surrogate-screened: deterministic for a seed, every reported best error
agrees within 1e-5 with the FP64 kernel and with the FP64 CPU oracle, and
the best never rises."""
import numpy as np
import pytest

import grn
import grn_oracle as ora
import evolve

pytestmark = pytest.mark.gpu


def check_best_with_oracle(prob, run):
    """The run's reported best error against the oracle's evaluation of its
    best model (same time step): within 1e-5 relative."""
    p = prob.params
    op = ora.params(p.dim_x, p.dim_y, p.n_input_morphogens, p.n_target_morphogens, p.num_sim_steps, run.best_dt,
                    fitness_kind=p.fitness_kind)
    want = ora.evaluate(op, prob.init, prob.target, run.best_model.genes, run.best_model.links)[0]["fitness"]
    assert abs(run.best_error - want) <= 1e-5 * want, (run.best_error, want)
    assert abs(run.best_exact - want) <= 1e-10 * want


def test_plain_and_screened_evolution_are_verified_and_deterministic():
    print("\n[gpu] evolution runs", flush=True)
    ctx = grn.Context(0)
    prob = evolve.scaled_config4(seed=2, n_islands=4, island_size=128, steps=400)
    evolve.make_target(ctx, prob)
    import os
    import tempfile
    import store
    path = os.path.join(tempfile.mkdtemp(), "evolve.db")
    with store.Store(path) as results:
        a = evolve.evolve(ctx, prob, False, max_generations=12, threshold=0.0, results=results, name="plain")
    b = evolve.evolve(ctx, prob, False, max_generations=12, threshold=0.0)
    with store.Store(path) as results:  # the run's record reads back
        gens = results.generations(1)
        assert [g["min_fit"] for g in gens] == a.best
        assert results.pareto_front(1)[0][1] == a.best[-1]
    assert a.best == b.best and a.verified and b.verified
    assert all(x >= y for x, y in zip(a.best, a.best[1:]))
    s = evolve.evolve(ctx, prob, True, max_generations=12, threshold=0.0, oversample=4, train_steps=5)
    assert s.verified and all(x >= y for x, y in zip(s.best, s.best[1:]))
    assert s.best[0] == a.best[0]  # the same initial population
    assert all(sc["n_candidates"] in (0, 4 * 512) for sc in s.screen)
    check_best_with_oracle(prob, a)
    check_best_with_oracle(prob, s)
    ctx.close()


def test_coarse_screen_equals_its_composition_and_ranks_candidates():
    """grn_search_run_generation_coarse keeps the children the public calls
    pick (candidates, a coarse simulation with scaled diffusion, screen),
    and the coarse errors rank the candidates like the full-size ones."""
    from scipy.stats import spearmanr
    print("\n[gpu] coarse screen", flush=True)
    ctx = grn.Context(0)
    prob = evolve.scaled_config4(seed=1, n_islands=2, island_size=64, steps=300)
    evolve.make_target(ctx, prob)
    ds = ctx.add_dataset(prob.params, prob.init, prob.target)
    ds_c, scale = evolve.coarse_dataset(ctx, prob, 8)
    assert scale == 1.0 / 16
    a = grn.Search(island_datasets=[ds, ds], **prob.search)
    b = grn.Search(island_datasets=[ds, ds], **prob.search)
    for s in (a, b):
        s.set_oversample(4)
        s.set_screen_policy(grn.SCREEN_ABSOLUTE)
    rhos = []
    for gen in range(3):
        a.run_generation_coarse(ctx, [ds_c, ds_c], scale)
        c = b.candidates()
        if c:
            dts = []
            for m in c:
                dmax = float(np.max(m.genes["dif"]))
                dts.append(min(prob.search["dt_scale"] / dmax, prob.search["dt_cap"]) if dmax > 0
                           else prob.search["dt_cap"])
            cm = []
            for m in c:
                g = m.genes.copy()
                g["dif"] = g["dif"] * scale
                cm.append(grn.Model(g, m.links))
            coarse = ctx.evaluate_batch(cm, datasets=[ds_c] * len(c), dt=dts)["raw_error"]
            fine = ctx.evaluate_batch(c, datasets=[ds] * len(c), dt=dts)["raw_error"]
            rhos.append(spearmanr(coarse, fine).correlation)
            b.screen(np.where(np.isnan(coarse), np.inf, np.log(np.maximum(coarse, 1e-300))))
        b.run_generation(ctx)
        pa, pb = a.pending(), b.pending()
        assert [m.to_text(17) for m in pa[0]] == [m.to_text(17) for m in pb[0]], gen
    print(f"[gpu] coarse/fine Spearman per generation: {rhos}", flush=True)
    assert len(rhos) == 2 and min(rhos) > 0.6
    a.close()
    b.close()
    ctx.close()


def test_coarse_screened_evolution_is_verified_and_deterministic():
    print("\n[gpu] coarse-screened evolution", flush=True)
    ctx = grn.Context(0)
    prob = evolve.scaled_config4(seed=2, n_islands=4, island_size=128, steps=400)
    evolve.make_target(ctx, prob)
    kw = dict(max_generations=12, threshold=0.0, oversample=4, screen="coarse", coarse_side=8,
              screen_policy=grn.SCREEN_ABSOLUTE)
    a = evolve.evolve(ctx, prob, True, **kw)
    b = evolve.evolve(ctx, prob, True, **kw)
    assert a.best == b.best and a.verified and b.verified
    assert all(x >= y for x, y in zip(a.best, a.best[1:]))
    p = evolve.evolve(ctx, prob, False, max_generations=1, threshold=0.0)
    assert a.best[0] == p.best[0]  # the same initial population, simulated unscreened
    assert all(x["coarse_ms"] >= 0 for x in a.screen)
    with pytest.raises(ValueError):
        evolve.evolve(ctx, prob, True, max_generations=1, threshold=0.0, screen="coarse", coarse_side=12)
    ctx.close()


def test_coarse_to_fine_generations_equal_their_composition():
    """grn_search_run_generation_scaled advances on the errors the public
    calls give (pending models with scaled diffusion on the coarse dataset,
    their own time steps and maxErrors), and after grn_search_rescore the
    members are simulated at full size as an initial population is."""
    print("\n[gpu] coarse-to-fine generations", flush=True)
    ctx = grn.Context(0)
    prob = evolve.scaled_config4(seed=3, n_islands=2, island_size=64, steps=300)
    evolve.make_target(ctx, prob)
    ds = ctx.add_dataset(prob.params, prob.init, prob.target)
    ds_c, scale = evolve.coarse_dataset(ctx, prob, 8)
    a = grn.Search(island_datasets=[ds, ds], **prob.search)
    b = grn.Search(island_datasets=[ds, ds], **prob.search)
    for gen in range(4):
        fine = gen == 3
        if gen == 3:
            a.rescore()
            b.rescore()
        if fine:
            ga = a.run_generation(ctx)
        else:
            ga = a.run_generation_scaled(ctx, [ds_c, ds_c], scale)
        models, dss, me, dts = b.pending()
        if fine:
            assert np.all(me == 1000.0) and len(models) == 128
            errs = ctx.evaluate_batch(models, datasets=dss, max_errors=me, dt=dts)["fitness"]
        else:
            cm = []
            for m in models:
                g = m.genes.copy()
                g["dif"] = g["dif"] * scale
                cm.append(grn.Model(g, m.links))
            errs = ctx.evaluate_batch(cm, datasets=[ds_c] * len(cm), max_errors=me, dt=dts)["fitness"]
        gb = b.advance(errs)
        assert ga == gb, gen
        pa, pb = a.pending(), b.pending()
        assert [m.to_text(17) for m in pa[0]] == [m.to_text(17) for m in pb[0]], gen
        assert np.array_equal(pa[2], pb[2])
    a.close()
    b.close()
    ctx.close()


def test_coarse_to_fine_evolution_is_verified_and_deterministic():
    print("\n[gpu] coarse-to-fine evolution", flush=True)
    ctx = grn.Context(0)
    prob = evolve.scaled_config4(seed=2, n_islands=4, island_size=128, steps=400)
    evolve.make_target(ctx, prob)
    for screened, kw, phases in ((False, {}, 6), (True, dict(oversample=4, screen="coarse"), 6),
                                 (False, {}, [(8, 3), (16, 3)])):
        runs = [evolve.evolve(ctx, prob, screened, max_generations=14, threshold=0.0, coarse_side=8,
                              coarse_generations=phases, **kw) for _ in range(2)]
        a, b = runs
        assert a.best == b.best and a.verified and b.verified
        assert a.best[:6] == [float("inf")] * 6 and np.isfinite(a.best[6])
        assert all(x >= y for x, y in zip(a.best[6:], a.best[7:]))
        assert [x["coarse_best"] for x in a.screen[:6]] == [x["coarse_best"] for x in b.screen[:6]]
        check_best_with_oracle(prob, a)
    with pytest.raises(ValueError):
        evolve.evolve(ctx, prob, False, max_generations=4, threshold=0.0, coarse_generations=4)
    ctx.close()
