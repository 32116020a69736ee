"""Time-to-fitness: the island search run to a fitness target, plain (every
child simulated) and surrogate-screened (k candidates per child, the best
predicted one simulated), on the same problem and seeds.

The search loop is Search::calcParalEvolution (Src/Search/search.cpp:101-216)
with every island advanced in lockstep, one batched evaluation per
generation; it stops when the best simulated error reaches the threshold
(the reference stops on a perfect model, search.cpp:159-167) or after
max_generations.  Every error the search keeps comes from the simulator.
"""
from __future__ import annotations

import math
import time
from dataclasses import dataclass, field

import numpy as np

import grn
import workloads
from grn import FITNESS_MSE, GENE_DTYPE, LINK_DTYPE, Model, SimParams


def hidden_blob_gate() -> Model:
    """Hidden 10-gene GRN with four inputs: genes 0, 1 hold an x- and a
    y-gradient, genes 2, 3 two Gaussian blobs (no production, no decay).
    Gene 5 is on in blob A's core, gene 6 right of x = 0.5, gene 8 above
    y = 0.5, gene 7 in blob B's core or above y = 0.5 (an OR pair), gene 9
    where 5 or 7 is; the target gene 4 is 9 AND NOT 6 -- the part of that
    region left of x = 0.5 -- and diffuses slightly."""
    genes = np.zeros(10, GENE_DTYPE)
    genes["label"] = np.arange(10)
    genes["lim"] = [0, 0, 0, 0, 1, 1, 1, 1, 1, 1]
    genes["deg"] = [0, 0, 0, 0, 1, 1, 1, 1, 1, 1]
    genes["dif"] = [0, 0, 0, 0, 0.02, 0, 0, 0, 0, 0]
    rows = [(2, 5, 0.4, 4.0, 0), (0, 6, 0.5, 6.0, 0), (3, 7, 0.4, 4.0, 0), (1, 8, 0.5, 6.0, 0),
            (5, 9, 0.5, 4.0, 0), (7, 9, 0.5, 4.0, 0), (9, 4, 0.4, 4.0, 1), (6, 4, 0.5, -4.0, 1),
            (8, 7, 0.9, 1.0, 0)]
    links = np.zeros(len(rows), LINK_DTYPE)
    for i, r in enumerate(rows):
        links[i] = r
    return Model(genes, links)


@dataclass
class Problem:
    name: str
    params: SimParams
    init: np.ndarray
    target: np.ndarray
    search: dict


def scaled_config4(seed: int = 0, n_islands: int = 8, island_size: int = 256, side: int = 32,
                   steps: int = 1000) -> Problem:
    """The full-evolution workload at a size one GPU runs in seconds per
    generation: four inputs (two blobs, two gradients), a hidden 10-gene GRN,
    genes in use capped at 10."""
    rng = np.random.default_rng(seed)
    blobs = [workloads.gaussian_blob(side, side, rng.uniform(0.2, 0.8) * side, rng.uniform(0.2, 0.8) * side,
                                     side / 6) for _ in range(2)]
    init = np.stack([workloads.gradient_x(side, side), workloads.gradient_y(side, side)] + blobs)[None]
    p = SimParams(side, side, 4, 1, steps, 1.0, fitness_kind=FITNESS_MSE)
    search = dict(n_islands=n_islands, island_size=island_size, n_morphogens=5, n_max_products=10,
                  n_max_links=30, init_products_min=5, init_products_max=10, migration=grn.MIGRATION_RING,
                  migration_period=10, migration_fraction=0.05, deg_lo=0.01, dif_prob=0.3, dif_hi=0.1,
                  k_lo=0.05, n_hi=4.0, dt_rule=1, dt_scale=0.2, dt_cap=1.0, seed=seed + 1)
    return Problem(f"{n_islands} islands x {island_size}, {side}x{side}, {steps} steps, hidden 10-gene GRN",
                   p, init, None, search)


def make_target(ctx, prob: Problem) -> np.ndarray:
    hidden = hidden_blob_gate()
    dummy = np.zeros((1, 1, prob.params.dim_y, prob.params.dim_x))
    ds = ctx.add_dataset(prob.params, prob.init, dummy)
    dt = workloads.dt_rule(hidden, prob.search["dt_scale"], prob.search["dt_cap"])
    f = ctx.simulate_fields(hidden, dataset=ds, dt=dt)
    prob.target = f[4:5][None].copy()
    return prob.target


@dataclass
class Run:
    screened: bool
    best: list = field(default_factory=list)      # best simulated error after each generation
    wall: list = field(default_factory=list)      # seconds since the start, after each generation
    cell_steps: float = 0.0
    screen: list = field(default_factory=list)    # per-generation screening stats

    def time_to(self, threshold: float):
        for b, t in zip(self.best, self.wall):
            if b <= threshold:
                return t
        return None


def coarse_dataset(ctx, prob: Problem, side: int):
    """The problem's experiments on a side x side grid (block means of the
    inputs and targets) and the factor that keeps the physical diffusion:
    (dataset id, dif_scale)."""
    W, H = prob.params.dim_x, prob.params.dim_y
    if W % side or H % side or W // side != H // side:
        raise ValueError(f"a {W}x{H} grid does not coarsen to {side}x{side}")
    f = W // side

    def down(a):
        return a.reshape(a.shape[:-2] + (a.shape[-2] // f, f, a.shape[-1] // f, f)).mean(axis=(-3, -1))

    p = prob.params
    pc = SimParams(side, side, p.n_input_morphogens, p.n_target_morphogens, p.num_sim_steps, p.dt,
                   min_conc=p.min_conc, min_conc_change=p.min_conc_change,
                   dist_error_threshold=p.dist_error_threshold, error_precision=p.error_precision,
                   fitness_kind=p.fitness_kind)
    return ctx.add_dataset(pc, down(prob.init), down(prob.target)), 1.0 / (f * f)


def evolve(ctx, prob: Problem, screened: bool, max_generations: int, threshold: float, oversample: int = 8,
           train_steps: int = 20, surrogate_batch: int = 4096, screen_fraction: float = 1.0,
           screen_policy: int = grn.SCREEN_ABSOLUTE, screen: str = "mlp", coarse_side: int = 8,
           pair_weight: float = 0.0, coarse_generations=0, time_limit: float = None, progress=None,
           results=None, name: str = "") -> Run:
    """Plain (screened False) or screened evolution.  screen "mlp": the
    learned surrogate scores the candidates; "coarse": a simulation of every
    candidate on a coarse_side x coarse_side grid does.  pair_weight > 0
    trains the surrogate on (child, slot parent) pairs as well
    (Surrogate.set_pair_weight).  coarse_generations > 0: a coarse-to-fine
    schedule -- the first coarse_generations generations simulate every
    child on the coarse_side grid only (Search.run_generation_scaled; their
    best counts as no simulator-verified error, so run.best holds inf), then
    the population is scored at full size (Search.rescore) and the run goes
    on as configured.  A list of (side, generations) phases instead runs
    the phases in turn, each on its own grid, the population scored again
    at every change of grid.  time_limit (seconds, optional): the run also
    ends after the generation that passes it.  results (optional):
    a store.Store that receives the run's parameters, every generation's
    statistics and the final Pareto front."""
    if screened and screen not in ("mlp", "coarse"):
        raise ValueError(f"unknown screen {screen!r}")
    phases = [(coarse_side, coarse_generations)] if isinstance(coarse_generations, int) else \
        [(int(a), int(b)) for a, b in coarse_generations]
    phases = [(side, n) for side, n in phases if n != 0]
    n_coarse = sum(n for _, n in phases)
    if any(n < 0 for _, n in phases) or n_coarse >= max_generations:
        raise ValueError("coarse generations must be >= 0 and fewer than max_generations in all")
    if prob.target is None:
        make_target(ctx, prob)
    ds = ctx.add_dataset(prob.params, prob.init, prob.target)
    if screened and screen == "coarse":
        ds_c, dif_scale = coarse_dataset(ctx, prob, coarse_side)
    # per coarse generation: (dataset, diffusion scale, last of its phase)
    schedule = []
    for side, n in phases:
        d, f = coarse_dataset(ctx, prob, side)
        schedule += [(d, f, i == n - 1) for i in range(n)]
    ctx.set_launch_mode(grn.LAUNCH_GRAPH)
    n_isl = prob.search["n_islands"]
    search = grn.Search(island_datasets=[ds] * n_isl, **prob.search)
    sur = None
    if screened:
        if screen == "mlp":
            sur = grn.Surrogate(0, n_morphogens=prob.search["n_morphogens"], batch=surrogate_batch,
                                seed=prob.search["seed"])
            if pair_weight > 0:
                sur.set_pair_weight(pair_weight)
        search.set_oversample(1 if schedule else oversample)
        search.set_screen_fraction(screen_fraction)
        search.set_screen_policy(screen_policy)
    run = Run(screened)
    if results is not None:
        sid = results.begin_search(name or prob.name, prob.search["seed"], results.add_sim_params(prob.params),
                                   results.add_search_params(search.params), n_isl)
    t0 = time.perf_counter()
    for gen in range(max_generations):
        if gen < n_coarse:
            d, f, last = schedule[gen]
            g = search.run_generation_scaled(ctx, [d] * n_isl, f)
            run.cell_steps += ctx.stats()["cell_steps"]
            run.best.append(math.inf)
            run.wall.append(time.perf_counter() - t0)
            run.screen.append({"coarse_best": g["min_fit"]})
            if last:  # the next grid scores the members again
                if gen + 1 == n_coarse and screened:  # to full size: k candidates per slot from now on
                    search.set_oversample(oversample)
                search.rescore()
            continue
        if screened and screen == "coarse":
            g, cms = search.run_generation_coarse(ctx, [ds_c] * n_isl, dif_scale)
            run.screen.append({"coarse_ms": cms})
        elif screened:
            g, sc = search.run_generation_screened(ctx, sur, train_steps=train_steps)
            run.screen.append(sc)
        else:
            g = search.run_generation(ctx)
        run.cell_steps += ctx.stats()["cell_steps"]
        run.best.append(g["min_fit"])
        run.wall.append(time.perf_counter() - t0)
        if results is not None:
            results.add_generation(sid, g, run.wall[-1])
        if progress and gen % 100 == 0:
            progress(f"[evolve] {'screened' if screened else 'plain'} gen {gen} best {g['min_fit']:.4g} "
                     f"t {run.wall[-1]:.1f}s")
        if g["min_fit"] <= threshold or g["done"]:  # target reached, or the search's own end (search.cpp:157-167)
            break
        if time_limit is not None and run.wall[-1] >= time_limit:
            break
    # the best model re-simulated by the FP64 kernel (the reference's double
    # precision throughout): the reported error must agree to 1e-5
    best, err = search.best()
    dmax = float(np.max(best.genes["dif"])) if len(best.genes) else 0.0
    dt = min(prob.search["dt_scale"] / dmax, prob.search["dt_cap"]) if dmax > 0 else prob.search["dt_cap"]
    exact = float(ctx.evaluate_batch([best], datasets=ds, dt=[dt], path=grn.PATH_EXACT)["fitness"][0])
    run.verified = bool(abs(exact - err) <= 1e-5 * abs(exact))
    run.best_model, run.best_error, run.best_exact, run.best_dt = best, err, exact, dt
    if results is not None:
        results.add_pareto_front(sid, search.pareto())
        results.end_search(sid)
    if sur is not None:
        sur.close()
    search.close()
    ctx.set_launch_mode(grn.LAUNCH_EAGER)
    return run
