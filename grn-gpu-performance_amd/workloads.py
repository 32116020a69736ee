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
