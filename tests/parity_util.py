"""Conditioning of a model's fitness, measured with the FP64 oracle itself.

The device kernels compute in FP32 (compensated state); their results
differ from the FP64 reference by roughly the unit roundoff of FP32
(6e-8) times the condition number of the model's fitness with respect to
its inputs.  Most models are well conditioned and the device meets the
1e-5 relative bar with orders of magnitude to spare; a model whose
dynamics is near a bifurcation or chaotic (steep Hill terms with explicit
steps) amplifies any rounding, and no two correct implementations agree on
it -- the reference's own GPU and CPU paths included.  The oracle
measures that directly: it re-evaluates the model with every parameter
(lim, deg, dif, K, n) scaled by 1 +- 1e-7 (two random sign patterns, a
change about the size of rounding the parameters to FP32) and reports the
largest relative change of the fitness.  A model whose fitness moves by
more than WELL (1e-6, a tenth of the parity bar) under that perturbation is
ill-conditioned at FP32 precision; for it the parity tests require the
device to stay within a fixed multiple of the oracle's own perturbation
spread instead of the flat 1e-5.
"""
import numpy as np

import grn
import grn_oracle as ora

PERTURBATION = 1e-7
WELL = 1e-6
FIT_RTOL = 1e-5
SPREAD_FACTOR = 20.0


def perturbed(m, rng, eps=PERTURBATION):
    g, l = m.genes.copy(), m.links.copy()
    for k in ("lim", "deg", "dif"):
        g[k] = g[k] * (1 + eps * rng.choice([-1.0, 1.0], len(g)))
    for k in ("dis_const", "hill_coef"):
        l[k] = l[k] * (1 + eps * rng.choice([-1.0, 1.0], len(l)))
    return grn.Model(g, l)


def oracle_with_conditioning(op, init, target, models, dt=None, n_threads=16, seed=0):
    """(oracle results, relative fitness spread under the perturbation) per model."""
    rng = np.random.default_rng(seed)
    want = ora.evaluate_batch(op, init, target, models, dt=dt, n_threads=n_threads)
    spread = np.zeros(len(models))
    for _ in range(2):
        p = ora.evaluate_batch(op, init, target, [perturbed(m, rng) for m in models], dt=dt, n_threads=n_threads)
        d = np.abs(p["raw_error"] - want["raw_error"])
        rel = np.where(want["raw_error"] != 0, d / np.where(want["raw_error"] != 0, np.abs(want["raw_error"]), 1.0),
                       np.where(d == 0, 0.0, np.inf))
        spread = np.maximum(spread, rel)
    return want, spread


def check_conditioned(got, want, spread):
    """Step counts exact; fitness within 1e-5 relative for every model that is
    well conditioned, within SPREAD_FACTOR times the oracle's own spread for
    the others; identical ranking among the well-conditioned models.
    Returns (well-conditioned mask, relative errors)."""
    assert np.array_equal(got["steps"], want["steps"])
    assert np.array_equal(got["n_products"], want["n_products"])
    w = want["raw_error"]
    diff = np.abs(got["raw_error"] - w)
    rel = np.where(w != 0, diff / np.where(w != 0, np.abs(w), 1.0), np.where(diff == 0, 0.0, np.inf))
    well = spread <= WELL
    bad = well & (rel > FIT_RTOL)
    assert not bad.any(), f"well-conditioned models beyond 1e-5: rel {rel[bad]}, spread {spread[bad]}"
    ill = ~well
    worse = ill & (rel > np.maximum(FIT_RTOL, SPREAD_FACTOR * spread))
    assert not worse.any(), f"ill-conditioned models beyond {SPREAD_FACTOR}x their spread: {rel[worse]} {spread[worse]}"
    assert same_ranking(got["raw_error"][well], w[well])
    return well, rel


def same_ranking(got, want, rtol=2 * FIT_RTOL):
    """Every pair of models the oracle tells apart by more than rtol
    (relative) is ordered the same way by the device; pairs the oracle
    scores equal within rtol (exact ties included, e.g. models that never
    express a target gene) carry no order."""
    got, want = np.asarray(got), np.asarray(want)
    gap = np.abs(want[:, None] - want[None, :])
    scale = np.maximum(np.abs(want[:, None]), np.abs(want[None, :]))
    apart = gap > rtol * scale
    return bool(np.all(np.sign(got[:, None] - got[None, :])[apart] == np.sign(want[:, None] - want[None, :])[apart]))
