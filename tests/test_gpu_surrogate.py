"""The surrogate's BF16-MFMA kernels against the FP64 oracle
(oracle/mlp_oracle.py), and what the trained network gains.

Tolerances: the oracle in bf16 mode rounds the same quantities the device
keeps in BF16 (inputs, weights, activations, backward signals), so device
and oracle differ by FP32 accumulation order and by rare BF16 rounding-
boundary flips: forward outputs within 1e-4 of the output scale for at
least 75 % of the rows and 1e-2 for all, gradients within 1e-2 relative L2 error per
parameter tensor (cosine > 0.9999).
Against the all-FP64 network (no BF16 anywhere) the gradient direction
agrees to cosine > 0.99.
"""
import numpy as np
import pytest

import grn
import mlp_oracle as mo
import workloads

pytestmark = pytest.mark.gpu

W, L = grn.MLP_WIDTH, 4


def rand_params(rng, scale=1.0):
    p = np.zeros(grn.n_surrogate_params(), np.float32)
    mat = W * W
    p[:L * mat] = rng.uniform(-1, 1, L * mat) * np.sqrt(6.0 / W) * scale
    p[L * mat:L * mat + L * W] = rng.normal(0, 0.1, L * W)
    p[L * mat + L * W:L * mat + L * W + W] = rng.uniform(-1, 1, W) * np.sqrt(3.0 / W)
    p[-1] = 0.3
    return p


def padded(F):
    X = np.zeros((len(F), W))
    X[:, :grn.MLP_FEATURES] = F
    return X


@pytest.fixture(scope="module")
def sur():
    print("\n[gpu] creating surrogate on device 0", flush=True)
    s = grn.Surrogate(0, n_morphogens=3, batch=1024, capacity=1 << 16, seed=7)
    yield s
    s.close()


def test_gemm_matches_fp64(sur):
    rng = np.random.default_rng(0)
    for M, N, K in [(128, 128, 32), (256, 384, 512), (512, 128, 96)]:
        print(f"[gpu] gemm {M}x{N}x{K}", flush=True)
        A = rng.normal(size=(M, K)).astype(np.float32)
        B = rng.normal(size=(N, K)).astype(np.float32)
        C = grn.mlp_gemm_bf16(mo.bf16_bits(A), mo.bf16_bits(B))
        ref = mo.bf16_round(A) @ mo.bf16_round(B).T
        bound = np.abs(mo.bf16_round(A)) @ np.abs(mo.bf16_round(B)).T
        assert np.all(np.abs(C - ref) <= 2e-6 * bound + 1e-30), np.max(np.abs(C - ref) / bound)


def test_forward_matches_oracle(sur):
    rng = np.random.default_rng(1)
    p = rand_params(rng)
    sur.set_params(p)
    assert np.array_equal(sur.get_params(), p)
    F = rng.uniform(-1, 1, (300, grn.MLP_FEATURES)).astype(np.float32)  # 300: padded to 384 rows
    st = sur.stats()
    got = sur.predict_features(F)  # ln error = y * sigma + mu
    y = (got - st["mu"]) / st["sigma"]
    want, _ = mo.forward(p, padded(F), W, L, bf16=True)
    scale = np.std(want) + 1e-12
    d = np.abs(y - want) / scale
    # most rows agree to FP32 accumulation; a row with an activation that
    # rounds to the other BF16 neighbour (each of its ~2000 roundings lands
    # within FP32 noise of a midpoint with probability ~5e-5) moves by ~2^-8
    # of one term
    assert np.median(d) <= 1e-5, np.median(d)
    assert np.mean(d > 1e-4) <= 0.25, np.mean(d > 1e-4)
    assert np.max(d) <= 1e-2, np.max(d)
    exact, _ = mo.forward(p, padded(F), W, L, bf16=False)
    assert np.corrcoef(y, exact)[0, 1] > 0.999


def test_gradient_matches_oracle(sur):
    rng = np.random.default_rng(2)
    p = rand_params(rng)
    sur.set_params(p)
    F = rng.uniform(-1, 1, (200, grn.MLP_FEATURES)).astype(np.float32)
    t = rng.normal(0, 1, 200)
    loss, g = sur.gradient(F, t)
    want_loss, want = mo.loss_and_grad(p, padded(F), t, W, L, bf16=True)
    assert loss == pytest.approx(want_loss, rel=1e-3)
    mat = W * W
    segs = [(l * mat, (l + 1) * mat) for l in range(L)]
    segs += [(L * mat + l * W, L * mat + (l + 1) * W) for l in range(L)]
    segs += [(L * mat + L * W, L * mat + L * W + W), (len(p) - 1, len(p))]
    for a, b in segs:
        ga, wa = g[a:b].astype(np.float64), want[a:b]
        rel = np.linalg.norm(ga - wa) / (np.linalg.norm(wa) + 1e-30)
        assert rel < 1e-2, (a, b, rel)
        if b - a > 1:
            assert ga @ wa / (np.linalg.norm(ga) * np.linalg.norm(wa)) > 0.9999
    _, exact = mo.loss_and_grad(p, padded(F), t, W, L, bf16=False)
    assert g @ exact / (np.linalg.norm(g) * np.linalg.norm(exact)) > 0.99


def test_pair_mode_gradient_matches_oracle():
    """Pair mode: rows 2i, 2i + 1 are (child, parent) and the loss adds the
    squared error of the predicted change; the device gradient against the
    oracle's, with the same bars as the plain loss."""
    rng = np.random.default_rng(12)
    s = grn.Surrogate(0, n_morphogens=3, batch=1024, capacity=1 << 12, seed=8)
    s.set_pair_weight(3.0)
    p = rand_params(rng)
    s.set_params(p)
    F = rng.uniform(-1, 1, (200, grn.MLP_FEATURES)).astype(np.float32)
    F[1::2] = F[0::2] + rng.normal(0, 0.05, F[1::2].shape).astype(np.float32)  # parents near their children
    t = rng.normal(0, 1, 200)
    loss, g = s.gradient(F, t)
    want_loss, want = mo.loss_and_grad(p, padded(F), t, W, L, bf16=True, pair_w=3.0)
    _, plain = mo.loss_and_grad(p, padded(F), t, W, L, bf16=True)
    assert loss == pytest.approx(want_loss, rel=1e-3)
    rel = np.linalg.norm(g - want) / np.linalg.norm(want)
    assert rel < 1e-2, rel
    assert np.linalg.norm(want - plain) > 10 * np.linalg.norm(g - want)  # the pair term is resolved
    with pytest.raises(grn.GrnError):
        s.gradient(F[:199], t[:199])  # pairs only
    s.close()
    s2 = grn.Surrogate(0, n_morphogens=3, batch=1024, capacity=1 << 12, seed=8)
    s2.add([workloads.config0(seed=1, n_models=1).models[0]], [0.5])
    with pytest.raises(grn.GrnError):
        s2.set_pair_weight(1.0)  # the store already holds single rows
    s2.close()


def test_adam_step_matches_oracle():
    rng = np.random.default_rng(3)
    s = grn.Surrogate(0, n_morphogens=3, batch=128, capacity=16, seed=3, lr=1e-3)
    p = rand_params(rng)
    s.set_params(p)
    m = workloads.random_model(rng, 5, 8)
    s.add([m], [0.25])  # one sample: every minibatch row is this sample
    F = grn.encode(m, 3)[None]
    # a single sample standardises to target 0 (sigma is floored at 1e-3)
    _, g = s.gradient(F, np.zeros(1))
    s.train(1)
    got = s.get_params().astype(np.float64)
    want, _, _ = mo.adam_step(p.astype(np.float64), 0, 0, g.astype(np.float64), 1, lr=1e-3)
    big = np.abs(g) > 1e-5  # where |g| >> eps the step is lr * sign(g)
    assert np.max(np.abs(got - want)[big]) < 2e-6
    # no step exceeds lr (plus the FP32 rounding of parameters below 1 in size)
    assert np.max(np.abs(got - p)) <= 1e-3 + 6e-8
    s.close()


def with_target_activator(rng, models, target_gene=2):
    """Gives every model an activator of its target gene (else the gene stays
    at zero and most errors tie)."""
    for m in models:
        l = m.links
        if np.any((l["regulated"] == target_gene) & (l["hill_coef"] > 0)):
            continue
        src = int(rng.integers(0, 2))
        keep = l[~((l["regulator"] == src) & (l["regulated"] == target_gene))]
        links = np.zeros(len(keep) + 1, grn.LINK_DTYPE)
        links[:-1] = keep
        links[-1] = (src, target_gene, rng.uniform(0.05, 1.0), rng.uniform(1, 4), 0)
        m.links = links
    return models


def simulated_errors(ctx, n_models, seed, side=32, steps=300):
    rng = np.random.default_rng(seed)
    p = grn.SimParams(side, side, 2, 1, steps, 0.05, fitness_kind=grn.FITNESS_MSE)
    init = np.stack([workloads.gradient_x(side, side), workloads.gradient_y(side, side)])[None]
    target = np.full((1, 1, side, side), 0.3)
    ds = ctx.add_dataset(p, init, target)
    models = with_target_activator(rng, workloads._population(rng, n_models, (3, 8), (2, 20), 3))
    res = ctx.evaluate_batch(models, datasets=ds)
    return models, res["raw_error"]


def test_training_learns_simulated_errors():
    print("\n[gpu] simulating training data", flush=True)
    ctx = grn.Context(0)
    models, err = simulated_errors(ctx, 6144, seed=11)
    s = grn.Surrogate(0, n_morphogens=3, batch=1024, capacity=1 << 14, seed=5)
    train, test = slice(0, 5120), slice(5120, 6144)
    s.add(models[train], err[train])
    before = s.predict(models[test])
    losses = [s.train(100) for _ in range(4)]
    print(f"[gpu] losses {losses}", flush=True)
    after = s.predict(models[test])
    from scipy.stats import spearmanr
    rho0 = spearmanr(before, np.log(err[test])).correlation
    rho = spearmanr(after, np.log(err[test])).correlation
    print(f"[gpu] held-out spearman before {rho0:.3f} after {rho:.3f}", flush=True)
    assert losses[-1] < losses[0]
    # random GRNs are only modestly predictable from their parameters (a
    # gradient-boosted tree ensemble on the same features reaches about 0.3)
    assert rho > 0.2 and rho > rho0 + 0.1
    st = s.stats()
    assert st["train_steps"] == 400 and st["samples"] == 5120
    s.close()
    ctx.close()


def search_problem(ctx, seed=4, side=32, steps=300):
    """A blob to grow from two gradients (errors are never all tied)."""
    rng = np.random.default_rng(seed)
    p = grn.SimParams(side, side, 2, 1, steps, 0.05, fitness_kind=grn.FITNESS_MSE)
    init = np.stack([workloads.gradient_x(side, side), workloads.gradient_y(side, side)])[None]
    target = workloads.gaussian_blob(side, side, rng.uniform(8, 24), rng.uniform(8, 24), 6.0)[None, None]
    return ctx.add_dataset(p, init, target)


SEARCH = dict(n_morphogens=3, n_max_products=8, n_max_links=16, init_products_min=3, init_products_max=8,
              migration=grn.MIGRATION_RING, migration_period=5, dif_hi=0.1)


def test_screening_beats_random_choice():
    """Ablation: among the same k = 8 candidates per offspring slot, the one
    the trained surrogate picks is better (by exact simulation) than a
    random pick."""
    ctx = grn.Context(0)
    ds = search_problem(ctx)
    s = grn.Search(island_datasets=[ds, ds], n_islands=2, island_size=1024, seed=8, **SEARCH)
    sur = grn.Surrogate(0, n_morphogens=3, batch=1024, capacity=1 << 15, seed=9)
    print("\n[gpu] generation 0 (simulated, trains the surrogate)", flush=True)
    models, dsi, me, dt = s.pending()
    res = ctx.evaluate_batch(models, datasets=dsi, max_errors=me, dt=dt)
    sur.add(models, res["raw_error"])
    sur.train(300)
    s.set_oversample(8)
    s.advance(res["fitness"])
    cands = s.candidates()
    n_slots = len(models)
    assert len(cands) == 8 * n_slots
    print(f"[gpu] simulating all {len(cands)} candidates", flush=True)
    exact = ctx.evaluate_batch(cands, datasets=ds)["raw_error"]
    pred = sur.predict(cands)
    # candidate order: island, repetition, slot
    per_island = 1024
    ranks, sel_err, all_err = [], [], []
    for isl in range(2):
        e = exact[isl * 8 * per_island:(isl + 1) * 8 * per_island].reshape(8, per_island)
        q = pred[isl * 8 * per_island:(isl + 1) * 8 * per_island].reshape(8, per_island)
        pick = np.argmin(q, axis=0)
        order = np.argsort(np.argsort(e, axis=0, kind="stable"), axis=0, kind="stable")
        ranks += list(order[pick, np.arange(per_island)])
        sel_err += list(np.log(e[pick, np.arange(per_island)]))
        all_err += list(np.log(e).ravel())
    mean_rank = float(np.mean(ranks))
    print(f"[gpu] mean exact rank of the surrogate's pick among 8: {mean_rank:.2f} (random: 3.5); "
          f"mean ln error picked {np.mean(sel_err):.3f} vs all {np.mean(all_err):.3f}", flush=True)
    # a random pick has mean rank 3.5 with standard error sqrt(63/12 / n);
    # demand a gain of five standard errors
    assert mean_rank < 3.5 - 5 * np.sqrt(63 / 12 / len(ranks))
    assert np.mean(sel_err) < np.mean(all_err)
    # screening then fills pending with exactly the picked candidates
    s.screen(pred)
    picked, _, _, _ = s.pending()
    assert len(picked) == n_slots
    sur.close()
    ctx.close()


def test_screened_generations_report_and_keep_simulated_errors():
    ctx = grn.Context(0)
    ds = search_problem(ctx, seed=5)
    s = grn.Search(island_datasets=[ds, ds], n_islands=2, island_size=128, seed=3, **SEARCH)
    sur = grn.Surrogate(0, n_morphogens=3, batch=512, capacity=1 << 14, seed=4)
    g, sc = s.run_generation_screened(ctx, sur, train_steps=50)
    assert sc["n_candidates"] == 0 and sc["n_simulated"] == 256 and sc["samples"] == 256
    s.set_oversample(4)
    g, sc = s.run_generation_screened(ctx, sur, train_steps=50)  # breeds the first candidates
    prev = g["min_fit"]
    for gen in range(3):
        print(f"[gpu] screened generation {gen}", flush=True)
        g, sc = s.run_generation_screened(ctx, sur, train_steps=50)
        assert sc["n_candidates"] == 4 * 256 and sc["n_simulated"] == 256
        assert -1.0 <= sc["spearman"] <= 1.0 and sc["mae"] >= 0
        assert g["min_fit"] <= prev
        prev = g["min_fit"]
    best, err = s.best()
    assert ctx.evaluate_batch([best], datasets=ds)["fitness"][0] == err
    sur.close()
    ctx.close()


def test_screened_generations_with_a_screen_fraction():
    """With screen fraction f only ceil(f n) slots per island are simulated;
    the best simulated error never rises and the reported best re-simulates
    to the same error."""
    ctx = grn.Context(0)
    ds = search_problem(ctx, seed=6)
    s = grn.Search(island_datasets=[ds, ds], n_islands=2, island_size=128, seed=4, **SEARCH)
    sur = grn.Surrogate(0, n_morphogens=3, batch=512, capacity=1 << 14, seed=5)
    s.set_oversample(4)
    s.set_screen_fraction(0.25)
    g, sc = s.run_generation_screened(ctx, sur, train_steps=50)  # initial population, all simulated
    assert sc["n_simulated"] == 256
    prev = g["min_fit"]
    for gen in range(4):
        print(f"[gpu] screened generation {gen}, fraction 0.25", flush=True)
        g, sc = s.run_generation_screened(ctx, sur, train_steps=50)
        assert sc["n_candidates"] == 4 * 256 and sc["n_simulated"] == 2 * 32
        assert g["n_evaluated"] == 64
        assert g["min_fit"] <= prev
        prev = g["min_fit"]
    best, err = s.best()
    assert ctx.evaluate_batch([best], datasets=ds)["fitness"][0] == err
    sur.close()
    ctx.close()


def test_screened_generations_with_the_better_half_policy():
    """SCREEN_HALF: the surrogate only rejects; generations run with every
    kept error simulated and the reported best re-simulating to itself."""
    ctx = grn.Context(0)
    ds = search_problem(ctx, seed=8)
    s = grn.Search(island_datasets=[ds, ds], n_islands=2, island_size=128, seed=9, **SEARCH)
    sur = grn.Surrogate(0, n_morphogens=3, batch=512, capacity=1 << 14, seed=3)
    s.set_oversample(4)
    s.set_screen_policy(grn.SCREEN_HALF)
    g, sc = s.run_generation_screened(ctx, sur, train_steps=20)
    prev = g["min_fit"]
    for gen in range(4):
        print(f"[gpu] better-half-screened generation {gen}", flush=True)
        g, sc = s.run_generation_screened(ctx, sur, train_steps=20)
        assert sc["n_simulated"] == 256
        assert g["min_fit"] <= prev
        prev = g["min_fit"]
    best, err = s.best()
    assert ctx.evaluate_batch([best], datasets=ds)["fitness"][0] == err
    sur.close()
    ctx.close()


def test_screened_generations_with_the_delta_policy():
    """SCREEN_DELTA: the surrogate learns the change of ln error from the slot
    parent (no samples from the unparented initial population); generations
    run, the best error never rises and re-simulates to the same value."""
    ctx = grn.Context(0)
    ds = search_problem(ctx, seed=7)
    s = grn.Search(island_datasets=[ds, ds], n_islands=2, island_size=128, seed=6, **SEARCH)
    sur = grn.Surrogate(0, n_morphogens=3, batch=512, capacity=1 << 14, seed=2)
    s.set_oversample(4)
    s.set_screen_policy(grn.SCREEN_DELTA)
    g, sc = s.run_generation_screened(ctx, sur, train_steps=20)
    assert sc["n_simulated"] == 256 and sc["samples"] == 0  # no parents yet
    g, sc = s.run_generation_screened(ctx, sur, train_steps=20)  # children of the initial population
    assert sc["samples"] > 0
    prev = g["min_fit"]
    for gen in range(3):
        print(f"[gpu] delta-screened generation {gen}", flush=True)
        g, sc = s.run_generation_screened(ctx, sur, train_steps=20)
        assert sc["n_candidates"] == 4 * 256 and sc["n_simulated"] == 256
        assert -1.0 <= sc["spearman"] <= 1.0 and np.isfinite(sc["mae"])
        assert g["min_fit"] <= prev
        prev = g["min_fit"]
    best, err = s.best()
    assert ctx.evaluate_batch([best], datasets=ds)["fitness"][0] == err
    with pytest.raises(grn.GrnError):
        s.set_screen_policy(7)
    sur.close()
    ctx.close()


def test_screened_generations_in_pair_mode():
    """Pair mode in a screened search: the initial population is stored as
    self-pairs, every later child beside its slot parent (two rows per
    simulated child); the best error never rises and re-simulates to itself."""
    ctx = grn.Context(0)
    ds = search_problem(ctx, seed=10)
    s = grn.Search(island_datasets=[ds, ds], n_islands=2, island_size=128, seed=11, **SEARCH)
    sur = grn.Surrogate(0, n_morphogens=3, batch=512, capacity=1 << 14, seed=6)
    sur.set_pair_weight(4.0)
    s.set_oversample(4)
    g, sc = s.run_generation_screened(ctx, sur, train_steps=20)
    assert sc["n_simulated"] == 256 and sc["samples"] == 2 * 256
    prev, samples = g["min_fit"], sc["samples"]
    for gen in range(4):
        print(f"[gpu] pair-mode screened generation {gen}", flush=True)
        g, sc = s.run_generation_screened(ctx, sur, train_steps=20)
        assert sc["n_simulated"] == 256
        assert sc["samples"] == samples + 2 * 256
        samples = sc["samples"]
        assert np.isfinite(sc["loss"]) and -1.0 <= sc["spearman"] <= 1.0
        assert g["min_fit"] <= prev
        prev = g["min_fit"]
    best, err = s.best()
    assert ctx.evaluate_batch([best], datasets=ds)["fitness"][0] == err
    sur.close()
    ctx.close()
