"""The evolutionary search that calls the path (grn_search_* of include/grn.h).

The reference draws its random numbers from an unseeded std::mt19937_64
(Src/Common/mathalgo.h:222-230), so its runs cannot be replayed and there are
no golden search trajectories.  These tests pin what the reference's code
fixes regardless of the random stream:

* the initial population of SearchAlgoDetCrowd::calcInitialPopulation
  (searchalgodetcrowd.cpp:54-61, model.cpp:285-315, modelprod.cpp:14-24,61-70);
* the coherence the reference asserts in debug builds after cross, mutate and
  createRandom (model.cpp:67-82,302-311,536-546): every link joins products
  of its model; plus one link per (regulator, regulated) pair
  (addOrReplaceRandomLink, model.cpp:449-453), morphogen products never
  removed (model.cpp:509,557-558) and the in-use caps of removeExcess
  (model.cpp:550-574);
* deterministic crowding: a child's maxError is its own parent's error
  (individual.cpp:19-25) and replaces that parent only when not worse
  (searchalgodetcrowd.cpp:119,128), so no island slot ever gets worse and
  the population's best and mean never rise;
* the 1 % duplicate-product rate of Model::mutate (model.cpp:505-506);
* pairwise full-mix migration (search.cpp:263-311) and the ring variant;
* a search driven end to end by the FP64 oracle as its evaluator.
Everything here runs on the CPU: the errors come from a deterministic fake
fitness or from the oracle.
"""
import numpy as np
import pytest

import grn
import grn_oracle as ora


def fake_error(m, dataset=0):
    """A deterministic error that depends on every parameter of the model."""
    g, l = m.genes, m.links
    v = abs(float(np.sum(g["lim"] * (1 + g["label"] % 7))) - 1.7) + 0.01 * len(l)
    v += 0.1 * abs(float(np.sum(l["dis_const"] * l["hill_coef"])) - 0.3)
    return 1.0 + (v % 1.0) * 0.999 + 1000.0 * dataset


def drive(search, gens, fitness=fake_error):
    """Runs gens generations with fitness as evaluator; returns (stats, pending) per generation."""
    hist = []
    for _ in range(gens):
        models, ds, me, dt = search.pending()
        errs = np.array([fitness(m, d) for m, d in zip(models, ds)])
        st = search.advance(errs)
        hist.append((st, search.pending(), errs))
    return hist


def check_coherent(m, n_morph, max_products, max_links):
    labels = [int(x) for x in m.genes["label"]]
    assert len(set(labels)) == len(labels), "duplicate product labels"
    assert set(range(n_morph)) <= set(labels), "a morphogen product was removed"
    pairs = set()
    for r, t in zip(m.links["regulator"], m.links["regulated"]):
        assert int(r) in labels and int(t) in labels, "link to a missing product"
        assert (int(r), int(t)) not in pairs, "two links with one (regulator, regulated) pair"
        pairs.add((int(r), int(t)))
    in_use = set(ora.labels_in_use(m.genes, m.links, n_morph))
    assert len(in_use) <= max_products
    links_in_use = sum(1 for r, t in pairs if r in in_use and t in in_use)
    assert links_in_use <= max_links


def test_initial_population():
    s = grn.Search(n_islands=3, island_size=33, n_morphogens=3, seed=7)
    models, ds, me, dt = s.pending()
    assert len(models) == 3 * 34  # island size rounded up to even (searchalgodetcrowd.cpp:33-35)
    assert np.all(me == 1000.0)   # Individual without parents (individual.cpp:23-24)
    assert np.all(ds == 0) and np.all(dt == 0)
    n_links = []
    for m in models:
        assert list(m.genes["label"]) == [0, 1, 2]
        assert list(m.genes["init"]) == [-1.0, -2.0, -3.0]  # zeroInit tags (modelprod.cpp:61-70)
        for k in ("lim", "deg", "dif"):
            assert np.all((m.genes[k] >= 0) & (m.genes[k] < 1))
        assert np.all((m.links["dis_const"] >= 0) & (m.links["dis_const"] < 1))
        n = np.abs(m.links["hill_coef"])
        assert np.all((n >= 1) & (n < 5))
        # repressors are AND links (modellink.cpp:21)
        assert np.all(m.links["is_and"][m.links["hill_coef"] < 0] == 1)
        check_coherent(m, 3, 16, 96)
        n_links.append(len(m.links))
    # each product regulates randInt(1 + n) = 0..3 targets, duplicates replaced
    assert 0 <= min(n_links) and max(n_links) <= 9
    assert 2.0 < np.mean(n_links) < 4.0
    # about 10 % of products diffuse (modelprod.cpp:19-22)
    dif = np.concatenate([m.genes["dif"] for m in models])
    assert 0.04 < np.mean(dif > 0) < 0.18


def test_replay_is_deterministic():
    a = grn.Search(n_islands=2, island_size=40, seed=11)
    b = grn.Search(n_islands=2, island_size=40, seed=11)
    c = grn.Search(n_islands=2, island_size=40, seed=12)
    ha, hb, hc = drive(a, 6), drive(b, 6), drive(c, 6)
    for (sa, pa, _), (sb, pb, _) in zip(ha, hb):
        assert sa == sb
        assert [m.to_text(17) for m in pa[0]] == [m.to_text(17) for m in pb[0]]
        assert np.array_equal(pa[2], pb[2])
    assert [m.to_text(17) for m in ha[-1][1][0]] != [m.to_text(17) for m in hc[-1][1][0]]


@pytest.mark.parametrize("migration", [grn.MIGRATION_PAIRWISE, grn.MIGRATION_RING, grn.MIGRATION_NONE])
def test_invariants_over_generations(migration):
    s = grn.Search(n_islands=4, island_size=48, n_morphogens=3, n_max_products=6, n_max_links=8,
                   migration=migration, migration_period=3, seed=3)
    submitted = set()
    prev = None
    for st, (models, ds, me, dt), errs in drive(s, 30):
        submitted.update(errs.tolist())
        for m in models:
            check_coherent(m, 3, 6, 8)
        # a child's maxError is an error its parent was given
        assert set(me.tolist()) <= submitted
        assert st["n_evaluated"] == len(errs)
        assert st["min_fit"] <= st["mean_fit"] <= st["max_fit"]
        if prev is not None:
            # no slot gets worse; migration only moves individuals
            assert st["min_fit"] <= prev["min_fit"]
            assert st["mean_fit"] <= prev["mean_fit"] + 1e-12
        assert st["pareto_size"] >= 1
        prev = st
    assert prev["generation"] == 30
    best, err = s.best()
    assert err == prev["min_fit"]
    assert fake_error(best) == err


def test_duplicate_product_rate():
    # with no crossover every child is its parent mutated once; a child of a
    # 3-product parent holds a new label (> 10, Model::createNewLabel) with
    # probability 1 - 0.99^3 ~= 2.97 %
    s = grn.Search(n_islands=1, island_size=8192, n_morphogens=3, crossover_rate=0.0, seed=5)
    drive(s, 1)
    models = s.pending()[0]
    frac = np.mean([np.any(m.genes["label"] > 10) for m in models])
    assert 0.022 < frac < 0.038
    assert all(np.all((m.genes["label"] < 3) | (m.genes["label"] > 10)) for m in models)


def _cross_island_children(pending, island_of_error):
    models, ds, me, dt = pending
    src = np.array([island_of_error(e) for e in me])
    return ds, src


def test_ring_migration_moves_the_best_fraction():
    n_isl, size, frac = 4, 40, 0.1
    s = grn.Search(island_datasets=list(range(n_isl)), n_islands=n_isl, island_size=size,
                   migration=grn.MIGRATION_RING, migration_period=2, migration_fraction=frac, seed=9)
    hist = drive(s, 2)
    # before the first migration every child's parent lived on its island
    ds, src = _cross_island_children(hist[0][1], lambda e: int(e // 1000))
    assert np.all(ds == src)
    # after it, island i+1 holds exactly round(frac * size) children of island i's best
    ds, src = _cross_island_children(hist[1][1], lambda e: int(e // 1000))
    k = round(frac * size)
    for i in range(n_isl):
        dst = (i + 1) % n_isl
        assert np.sum((ds == dst) & (src == i)) == k
        assert np.sum(ds == dst) == size
    assert np.sum(ds != src) == n_isl * k


def test_pairwise_migration_mixes_partner_islands():
    n_isl, size = 4, 64
    s = grn.Search(island_datasets=list(range(n_isl)), n_islands=n_isl, island_size=size,
                   migration=grn.MIGRATION_PAIRWISE, migration_period=2, seed=21)
    hist = drive(s, 2)
    ds, src = _cross_island_children(hist[0][1], lambda e: int(e // 1000))
    assert np.all(ds == src)
    ds, src = _cross_island_children(hist[1][1], lambda e: int(e // 1000))
    for i in range(n_isl):
        assert np.sum(ds == i) == size
        partners = set(src[ds == i].tolist())
        assert len(partners) == 2 and i in partners  # own island plus one partner
        other = (partners - {i}).pop()
        # the partner's children came from the same pair
        assert set(src[ds == other].tolist()) == {i, other}
        # full mix: each island keeps about half of its own
        assert 0.25 * size < np.sum((ds == i) & (src == i)) < 0.75 * size


def test_no_migration_keeps_islands_apart():
    s = grn.Search(island_datasets=[0, 1, 2], n_islands=3, island_size=32,
                   migration=grn.MIGRATION_NONE, migration_period=1, seed=2)
    for _, pend, _ in drive(s, 5):
        ds, src = _cross_island_children(pend, lambda e: int(e // 1000))
        assert np.all(ds == src)


def test_dt_rule():
    s = grn.Search(n_islands=1, island_size=256, dt_rule=1, dt_scale=0.2, dt_cap=1.0, dif_prob=0.5,
                   seed=4)
    models, _, _, dt = s.pending()
    for m, d in zip(models, dt):
        dmax = float(np.max(m.genes["dif"]))
        want = min(0.2 / dmax, 1.0) if dmax > 0 else 1.0
        assert d == pytest.approx(want, rel=1e-15)


@pytest.mark.parametrize("bad", [dict(n_islands=0), dict(island_size=1), dict(n_max_products=17),
                                 dict(n_max_links=97), dict(n_max_products=2, n_morphogens=3),
                                 dict(migration=5), dict(crossover_rate=1.5),
                                 dict(init_products_min=5, init_products_max=4),
                                 dict(dt_rule=1, dt_cap=0.0)])
def test_invalid_params_are_rejected(bad):
    with pytest.raises(grn.GrnError):
        grn.Search(**bad)


def test_advance_rejects_bad_errors():
    s = grn.Search(n_islands=1, island_size=4)
    with pytest.raises(grn.GrnError):
        s.advance(np.array([0.1, -1.0, 0.2, 0.3]))
    with pytest.raises(grn.GrnError):
        s.advance(np.array([0.1, np.nan, 0.2, 0.3]))
    with pytest.raises(grn.GrnError):
        s.best()  # nothing evaluated yet


def test_search_driven_by_the_oracle():
    # a tiny morphogen problem: 2 inputs (x and y gradients), 1 target
    W = H = 8
    x = np.linspace(0, 1, W)
    init = np.zeros((1, 2, H, W))
    init[0, 0] = x[None, :]
    init[0, 1] = x[:, None]
    target = np.zeros((1, 1, H, W))
    target[0, 0] = 0.5 * x[None, :] * x[:, None]
    p = ora.params(W, H, 2, 1, 60, 0.05, fitness_kind=1)
    s = grn.Search(n_islands=2, island_size=32, n_morphogens=3, n_max_products=6, n_max_links=10,
                   migration_period=2, seed=1)
    prev = None
    for _ in range(4):
        models, ds, me, dt = s.pending()
        res = ora.evaluate_batch(p, init, target, models, max_errors=me, n_threads=4)
        st = s.advance(res["fitness"])
        if prev is not None:
            assert st["min_fit"] <= prev["min_fit"]
        prev = st
    best, err = s.best()
    r, _ = ora.evaluate(p, init, target, best.genes, best.links)
    assert r["fitness"] == err


def test_screen_fraction_keeps_the_slots_that_promise_most():
    """grn_search_set_screen_fraction: per island, only the ceil(f n) slots
    whose kept candidate most undercuts the parent's recorded ln raw error are
    simulated; the others keep their parents."""
    n_isl, size, k, f = 2, 40, 3, 0.25
    s = grn.Search(n_islands=n_isl, island_size=size, seed=5)
    models, ds, me, dt = s.pending()
    errs = np.array([fake_error(m) for m in models])
    s.record_raw(np.ones(len(models)))  # every parent: ln raw error 0
    s.set_oversample(k)
    s.set_screen_fraction(f)
    s.advance(errs)
    cands = s.candidates()
    assert len(cands) == n_isl * k * size
    # one score per slot, shared by the slot's k candidates (order: island,
    # repetition, slot), so the kept candidate's score is the slot's value
    rng = np.random.default_rng(1)
    v = rng.permutation(n_isl * size).reshape(n_isl, size).astype(float)
    scores = np.repeat(v[:, None, :], k, axis=1).ravel()
    s.screen(scores)
    picked = [m.to_text(17) for m in s.pending()[0]]
    keep = int(np.ceil(f * size))
    assert len(picked) == n_isl * keep
    texts = np.array([m.to_text(17) for m in cands], dtype=object).reshape(n_isl, k, size)
    at = 0
    for isl in range(n_isl):
        slots = np.sort(np.argsort(v[isl], kind="stable")[:keep])  # pending keeps slot order
        for slot in slots:
            assert picked[at] in set(texts[isl, :, slot])
            at += 1
    # the unsimulated slots cannot replace their parents: the search goes on
    st = s.advance(np.array([fake_error(m) for m in s.pending()[0]]))
    assert st["n_evaluated"] == n_isl * keep
    assert len(s.candidates()) == n_isl * k * size


def test_screen_fraction_without_recorded_errors_and_bad_values():
    s = grn.Search(n_islands=1, island_size=20, seed=2)
    s.set_oversample(2)
    s.set_screen_fraction(0.5)
    s.advance(np.array([fake_error(m) for m in s.pending()[0]]))
    s.screen(np.zeros(len(s.candidates())))
    assert len(s.pending()[0]) == 10  # no parent has a recorded error: any 10 slots
    for bad in (0.0, -0.5, 1.5, float("nan")):
        with pytest.raises(grn.GrnError):
            s.set_screen_fraction(bad)
    with pytest.raises(ValueError):
        s.record_raw(np.ones(3))


def test_screen_policy_better_half_draws_reproducibly():
    """GRN_SCREEN_HALF: each slot's child is one of its candidates, drawn from
    the better-scored half by a seeded hash (the same on every run), so the
    picks differ from the lowest-score picks in some slots."""
    def picks(policy, seed=4):
        s = grn.Search(n_islands=2, island_size=32, seed=seed)
        s.set_oversample(4)
        s.advance(np.array([fake_error(m) for m in s.pending()[0]]))
        s.set_screen_policy(policy)
        cands = [m.to_text(17) for m in s.candidates()]
        rng = np.random.default_rng(7)
        s.screen(rng.permutation(len(cands)).astype(float))
        return cands, [m.to_text(17) for m in s.pending()[0]]
    cands, half = picks(grn.SCREEN_HALF)
    _, half2 = picks(grn.SCREEN_HALF)
    _, best = picks(grn.SCREEN_ABSOLUTE)
    assert half == half2
    assert set(half) <= set(cands) and len(half) == 64
    assert sum(a != b for a, b in zip(half, best)) > 0


def test_coarse_generation_rejects_bad_arguments():
    # argument checks run before any device work (no context is needed);
    # -1 is GRN_E_INVALID
    import ctypes
    s = grn.Search(n_islands=2, island_size=4)
    lib = grn.load_library()
    ds = np.zeros(2, np.int32)
    assert lib.grn_search_run_generation_coarse(s._h, None, ds.ctypes.data, 0.25, grn.PATH_AUTO, None, None) \
        == -1
    assert lib.grn_search_run_generation_coarse(None, None, ds.ctypes.data, 0.25, grn.PATH_AUTO, None, None) \
        == -1
    fake_ctx = ctypes.c_void_p(8)  # never dereferenced: the scale check comes first
    for bad in (-1.0, float("inf"), float("nan")):
        assert lib.grn_search_run_generation_coarse(s._h, fake_ctx, ds.ctypes.data, bad, grn.PATH_AUTO,
                                                    None, None) == -1
    with pytest.raises(ValueError):
        s.run_generation_coarse(None, [0, 0, 0], 0.25)  # one coarse dataset per island
    # the same checks guard a generation simulated at another fidelity
    assert lib.grn_search_run_generation_scaled(s._h, None, ds.ctypes.data, 0.25, grn.PATH_AUTO, None) == -1
    assert lib.grn_search_run_generation_scaled(s._h, fake_ctx, None, 0.25, grn.PATH_AUTO, None) == -1
    for bad in (-1.0, float("inf"), float("nan")):
        assert lib.grn_search_run_generation_scaled(s._h, fake_ctx, ds.ctypes.data, bad, grn.PATH_AUTO,
                                                    None) == -1
    with pytest.raises(ValueError):
        s.run_generation_scaled(None, [0], 0.25)
    assert lib.grn_search_rescore(None) == -1
    s.close()


def test_rescore_scores_the_members_again():
    """grn_search_rescore makes the members pending as an initial population
    is: the next advance records their new errors without selection (the
    statistics and the front see only those), with no maxError cut, then
    breeding goes on from them."""
    s = grn.Search(n_islands=2, island_size=8, n_morphogens=3, seed=7, migration=grn.MIGRATION_NONE)
    hist = drive(s, 3)
    before = hist[-1][0]
    s.rescore()
    models, ds, me, dt = s.pending()
    assert len(models) == 16 and np.all(me == 1000.0)
    errs = np.array([fake_error(m) for m in models])
    # the pending models are the population: the same errors give the same statistics
    assert errs.min() == before["min_fit"] and errs.max() == before["max_fit"]
    assert abs(errs.mean() - before["mean_fit"]) <= 1e-12
    # scored again (say at another fidelity): every member survives with its new error
    new = errs + 5.0
    st = s.advance(new)
    assert st["generation"] == before["generation"] + 1 and st["n_evaluated"] == 16
    assert st["min_fit"] == new.min() and st["max_fit"] == new.max()
    assert all(e["error"] in set(new) for e in s.pareto())
    # the next children compete with the re-scored parents
    kids, _, me2, _ = s.pending()
    assert len(kids) == 16 and set(me2) <= set(new)
    st2 = s.advance(np.array([fake_error(m) + 5.0 for m in kids]))
    assert st2["max_fit"] <= new.max()  # crowding keeps the better of child and parent
    s.close()


def complexity_in_use(m, n_morph):
    """Model::calcComplexityInUse (model.cpp:590-611): 1 + (dif > 0) per
    product in use (modelprod.cpp:39-41), 1 per link between products in use."""
    used = set(int(x) for x in ora.labels_in_use(m.genes, m.links, n_morph))
    c = sum(1 + (d > 0) for lab, d in zip(m.genes["label"], m.genes["dif"]) if int(lab) in used)
    return c + sum(1 for r, t in zip(m.links["regulator"], m.links["regulated"]) if int(r) in used and int(t) in used)


def ref_front(front, fresh):
    """Search::recalculateParetoFront (search.cpp:348-377) on (error,
    complexity) pairs: sort by error then complexity, keep each entry less
    complex than the last kept one, and exact duplicates of it."""
    out = []
    for e, c in sorted(front + fresh):
        if not out or c < out[-1][1] or (c == out[-1][1] and e == out[-1][0]):
            out.append((e, c))
    return out


def coarse_error(m, dataset=0):
    """fake_error rounded to 0.05: many exact ties, so the front's duplicate
    rule is exercised."""
    return round(fake_error(m, dataset) * 20) / 20


def test_pareto_front_follows_the_reference_rule():
    """The front the library keeps equals the reference's rule applied to the
    initial population and every selected child (a child is selected when
    its error is at most its parent's, which is its maxError), generation by
    generation; each member's model reproduces its error and complexity."""
    n_morph = 3
    s = grn.Search(n_islands=3, island_size=24, n_morphogens=n_morph, migration_period=4, seed=21)
    front = []
    for gen in range(25):
        models, ds, me, dt = s.pending()
        errs = np.array([coarse_error(m, d) for m, d in zip(models, ds)])
        fresh = [(float(e), complexity_in_use(m, n_morph)) for m, e, mx in zip(models, errs, me)
                 if gen == 0 or e <= mx]
        front = ref_front(front, fresh)
        st = s.advance(errs)
        got = s.pareto()
        assert [(x["error"], x["complexity"]) for x in got] == front, gen
        assert st["pareto_size"] == len(front)
    assert any(a == b for a, b in zip(front, front[1:]))  # duplicates kept
    for x in got:
        assert coarse_error(x["model"]) == x["error"]
        assert complexity_in_use(x["model"], n_morph) == x["complexity"]
        assert 0 <= x["generation"] < 25 and 0 <= x["island"] < 3
    assert st["done"] == 0 and st["last_generation"] == -1 and st["best_perfect_complexity"] == -1


def test_perfect_model_moves_the_last_generation():
    """search.cpp:157-167: once the front's head has error 0 with a new
    lowest complexity, the run ends 250 generations later (never after
    nGenerations); the search reports done when it gets there."""
    def perfect_if_small(m, dataset=0):
        # error 0 for models with at most two links, else the fake error
        return 0.0 if len(m.links) <= 2 else fake_error(m, dataset)

    s = grn.Search(n_islands=2, island_size=16, n_morphogens=3, seed=5, n_generations=1000)
    st = s.advance(np.array([perfect_if_small(m) for m in s.pending()[0]]))
    head = s.pareto()[0]
    assert head["error"] == 0.0
    assert st["best_perfect_complexity"] == head["complexity"]
    assert st["last_generation"] == 250 and st["done"] == 0
    s2 = grn.Search(n_islands=2, island_size=16, n_morphogens=3, seed=5, n_generations=100)
    st2 = s2.advance(np.array([perfect_if_small(m) for m in s2.pending()[0]]))
    assert st2["last_generation"] == 100  # capped by nGenerations
    # with a limit of 3 generations the search is done after generation 3
    s3 = grn.Search(n_islands=1, island_size=8, n_morphogens=3, seed=2, n_generations=3)
    done = [st["done"] for st, _, _ in drive(s3, 5)]
    assert done == [0, 0, 0, 1, 1]
