"""Seeded synthetic GRN workloads: random models, input patterns, targets.

Parameter distributions follow the benchmark configurations; the random
construction follows the reference's rules where they are defined
(Src/Model/modellink.cpp:10-22: a repressor is always an AND link, an
activator is AND with probability 1/2; Src/Model/model.cpp:449-453: at most one
link per (regulator, regulated) pair).  Genes 0..n_in-1 are the input
morphogens, n_in..n_in+n_tgt-1 the target genes, the rest are free genes.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from grn import FITNESS_MSE, GENE_DTYPE, LINK_DTYPE, Model, SimParams


def random_model(rng: np.random.Generator, n_genes: int, n_links: int, *,
                 lim=(0.0, 1.0), deg=(0.01, 1.0), dif=(0.0, 0.1), K=(0.05, 1.0),
                 n=(1.0, 4.0)) -> Model:
    genes = np.zeros(n_genes, GENE_DTYPE)
    genes["label"] = np.arange(n_genes)
    genes["lim"] = rng.uniform(*lim, n_genes)
    genes["deg"] = rng.uniform(*deg, n_genes)
    genes["dif"] = rng.uniform(*dif, n_genes)
    n_links = int(min(n_links, n_genes * n_genes))
    pairs = rng.choice(n_genes * n_genes, size=n_links, replace=False)
    links = np.zeros(n_links, LINK_DTYPE)
    links["regulator"] = pairs // n_genes
    links["regulated"] = pairs % n_genes
    links["dis_const"] = rng.uniform(*K, n_links)
    act = rng.random(n_links) < 0.5
    hill = rng.uniform(*n, n_links)
    links["hill_coef"] = np.where(act, hill, -hill)
    links["is_and"] = np.where(act, rng.random(n_links) < 0.5, True).astype(np.int32)
    return Model(genes, links)


def gradient_x(W: int, H: int) -> np.ndarray:
    return np.broadcast_to(np.arange(W, dtype=np.float64) / max(W - 1, 1), (H, W)).copy()


def gradient_y(W: int, H: int) -> np.ndarray:
    return np.broadcast_to((np.arange(H, dtype=np.float64) / max(H - 1, 1))[:, None], (H, W)).copy()


def gaussian_blob(W: int, H: int, cx: float, cy: float, sigma: float) -> np.ndarray:
    x = np.arange(W, dtype=np.float64)[None, :]
    y = np.arange(H, dtype=np.float64)[:, None]
    return np.exp(-((x - cx) ** 2 + (y - cy) ** 2) / (2.0 * sigma * sigma))


@dataclass
class Workload:
    name: str
    params: SimParams
    init: np.ndarray         # [1][n_in][H][W]
    hidden: Model            # the GRN whose output is the target
    models: list             # population
    target: np.ndarray | None = None  # [1][n_tgt][H][W], filled by make_target


def ensure_activated(model: Model, target_gene: int, n_inputs: int, seed: int) -> Model:
    """A hidden network whose target gene is activated by no input gene may
    produce zero there everywhere (h = 0 without an activator, modelsimulator.cpp:
    120-138, and a self-activation from zero stays at zero), which makes every
    model's fitness against it trivial.  Such a network gets one OR
    activation of the target gene from an input gene that does not regulate
    it yet, drawn from its own stream (seed), so the workload's other draws
    are unchanged."""
    links = model.links
    into = links[links["regulated"] == target_gene]
    if np.any((into["hill_coef"] >= 0) & (into["regulator"] < n_inputs)):
        return model
    rng = np.random.default_rng([seed, 0x5EED])
    free = [i for i in range(n_inputs) if i not in set(int(r) for r in into["regulator"])]
    if not free:  # every input already represses it: turn the first such link around
        l = links.copy()
        k = int(np.where((l["regulated"] == target_gene) & (l["regulator"] < n_inputs))[0][0])
        l["hill_coef"][k] = abs(l["hill_coef"][k])
        return Model(model.genes.copy(), l)
    new = np.zeros(1, LINK_DTYPE)
    new[0] = (int(rng.choice(free)), target_gene, rng.uniform(0.05, 1.0), rng.uniform(1.0, 4.0), 0)
    l = np.zeros(len(links) + 1, LINK_DTYPE)
    l[:len(links)] = links
    l[len(links):] = new
    return Model(model.genes.copy(), l)


def _population(rng, n_models, g_range, l_range, n_struct, **kw):
    out = []
    for _ in range(n_models):
        G = max(int(rng.integers(g_range[0], g_range[1] + 1)), n_struct)
        L = int(rng.integers(l_range[0], l_range[1] + 1))
        out.append(random_model(rng, G, L, **kw))
    return out


def config0(seed: int = 0, n_models: int = 64) -> Workload:
    """One island of 64 random 3-gene models on a 32x32 grid, 1000 steps."""
    rng = np.random.default_rng(seed)
    W = H = 32
    p = SimParams(W, H, 2, 1, 1000, 0.01, fitness_kind=FITNESS_MSE)
    init = np.stack([gradient_x(W, H), gradient_y(W, H)])[None]
    hidden = random_model(rng, 3, int(rng.integers(2, 7)))
    models = _population(rng, n_models, (3, 3), (2, 6), 3)
    hidden = ensure_activated(hidden, 2, 2, seed)
    return Workload("config0: 64 models, 3 genes, 32x32, 1000 steps", p, init, hidden, models)


def config1(seed: int = 0, n_models: int = 4096) -> Workload:
    """4096 models with 3-8 genes and 2-24 links on a 64x64 grid, 2000 steps.

    Inputs are an x-gradient, a y-gradient and a Gaussian blob (sigma 8 cells)
    at a random centre, so four genes are structural (3 inputs + 1 target) and
    a drawn gene count of 3 is raised to 4.
    """
    rng = np.random.default_rng(seed)
    W = H = 64
    p = SimParams(W, H, 3, 1, 2000, 0.01, fitness_kind=FITNESS_MSE)
    cx, cy = rng.uniform(0, W - 1), rng.uniform(0, H - 1)
    init = np.stack([gradient_x(W, H), gradient_y(W, H), gaussian_blob(W, H, cx, cy, 8.0)])[None]
    hidden = random_model(rng, 6, int(rng.integers(6, 19)))
    models = _population(rng, n_models, (3, 8), (2, 24), 4)
    hidden = ensure_activated(hidden, 3, 3, seed)
    return Workload("config1: 4096 models, 3-8 genes, 2-24 links, 64x64, 2000 steps", p, init,
                    hidden, models)


def make_target(ctx, w: Workload) -> np.ndarray:
    """Simulates the hidden GRN on the device and uses its target genes'
    final field as the target pattern."""
    dummy = np.zeros((1, w.params.n_target_morphogens, w.params.dim_y, w.params.dim_x))
    ds = ctx.add_dataset(w.params, w.init, dummy)
    f = ctx.simulate_fields(w.hidden, dataset=ds)
    n_in, n_tgt = w.params.n_input_morphogens, w.params.n_target_morphogens
    w.target = f[n_in:n_in + n_tgt][None].copy()
    return w.target


def _link(r, t, K, n, is_and):
    return (r, t, K, n, int(is_and))


def hidden_stripe_spot() -> Model:
    """The hidden 8-gene GRN of the multi-island workload.

    Genes 0 and 1 hold an x- and a y-gradient (no production, no decay); genes 4..7 switch on where the
    gradient passes 0.3 or 0.6 (steep Hill activations); gene 2 is on where
    x lies between the two thresholds (a stripe: activated by 4 AND repressed
    by 5) and gene 3 where both x and y do (a spot: the same in both
    directions).  Genes 2 and 3 diffuse slightly, which blurs the edges.
    """
    genes = np.zeros(8, GENE_DTYPE)
    genes["label"] = np.arange(8)
    genes["lim"] = [0.0, 0.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0]
    genes["deg"] = [0.0, 0.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0]  # the input planes stay fixed
    genes["dif"] = [0.0, 0.0, 0.02, 0.02, 0.0, 0.0, 0.0, 0.0]
    links = np.array([
        _link(0, 4, 0.3, 6.0, False), _link(0, 5, 0.6, 6.0, False),
        _link(1, 6, 0.3, 6.0, False), _link(1, 7, 0.6, 6.0, False),
        _link(4, 2, 0.5, 4.0, True), _link(5, 2, 0.5, -4.0, True),
        _link(4, 3, 0.5, 4.0, True), _link(5, 3, 0.5, -4.0, True),
        _link(6, 3, 0.5, 4.0, True), _link(7, 3, 0.5, -4.0, True),
    ], dtype=LINK_DTYPE)
    return Model(genes, links)


def dt_rule(model: Model, scale: float = 0.2, cap: float = 1.0) -> float:
    """dt = scale * dx^2 / max(D) with dx = 1, at most cap (cap when no gene diffuses)."""
    d = float(np.max(model.genes["dif"])) if len(model.genes) else 0.0
    return min(scale / d, cap) if d > 0 else cap


SIDES = (32, 48, 64, 96, 128)


@dataclass
class IslandWorkload:
    name: str
    sides: list              # grid side of each island
    params: list             # SimParams per island
    init: list               # [1][2][s][s] per island
    target: list             # [1][2][s][s] per island (filled by make_island_targets)
    hidden: Model
    search: dict             # keyword arguments of grn.Search


def config2(seed: int = 0, n_islands: int = 16, island_size: int = 1024, steps: int = 3000,
            sides=None) -> IslandWorkload:
    """16 islands x 1024 models, 2-12 genes (4 structural), 1-40 links, grid
    side per island from {32, 48, 64, 96, 128} (every side used at least once
    when there are 5 or more islands), 3000 steps, dt = 0.2 dx^2 / max(D)
    capped at 1.  Targets: stripe and spot of the hidden 8-gene GRN."""
    rng = np.random.default_rng(seed)
    if sides is None:
        sides = list(SIDES[:min(n_islands, len(SIDES))])
        sides += [int(s) for s in rng.choice(SIDES, n_islands - len(sides))]
    params, init = [], []
    for s in sides:
        params.append(SimParams(s, s, 2, 2, steps, 1.0, fitness_kind=FITNESS_MSE))
        init.append(np.stack([gradient_x(s, s), gradient_y(s, s)])[None])
    search = dict(n_islands=n_islands, island_size=island_size, n_morphogens=4, n_max_products=12,
                  n_max_links=40, init_products_min=4, init_products_max=12,
                  migration=1, migration_period=10, migration_fraction=0.05,
                  lim_lo=0.0, lim_hi=1.0, deg_lo=0.01, deg_hi=1.0, dif_prob=0.5, dif_lo=0.0,
                  dif_hi=0.1, k_lo=0.05, k_hi=1.0, n_lo=1.0, n_hi=4.0,
                  dt_rule=1, dt_scale=0.2, dt_cap=1.0, seed=seed + 1)
    return IslandWorkload(f"config2: {n_islands} islands x {island_size}, sides {sorted(set(sides))}, "
                          f"{steps} steps", sides, params, init, [None] * n_islands,
                          hidden_stripe_spot(), search)


def make_island_targets(ctx, w: IslandWorkload) -> list:
    """Simulates the hidden GRN on every island's grid (at its own dt rule)."""
    dt = dt_rule(w.hidden, w.search["dt_scale"], w.search["dt_cap"])
    cache = {}
    for i, (p, init) in enumerate(zip(w.params, w.init)):
        if p.dim_x not in cache:
            dummy = np.zeros((1, 2, p.dim_y, p.dim_x))
            ds = ctx.add_dataset(p, init, dummy)
            f = ctx.simulate_fields(w.hidden, dataset=ds, dt=dt)
            cache[p.dim_x] = f[2:4][None].copy()
        w.target[i] = cache[p.dim_x]
    return w.target


def add_island_datasets(ctx, w: IslandWorkload) -> list:
    """Uploads one dataset per island; returns the dataset ids."""
    if any(t is None for t in w.target):
        make_island_targets(ctx, w)
    return [ctx.add_dataset(p, i, t) for p, i, t in zip(w.params, w.init, w.target)]
