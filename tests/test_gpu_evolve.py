"""Evolution to a fitness target through the batched path, plain and
surrogate-screened: deterministic for a seed, every reported best error
re-simulates to the same value, and the best never rises."""
import numpy as np
import pytest

import grn
import evolve

pytestmark = pytest.mark.gpu


def test_plain_and_screened_evolution_are_verified_and_deterministic():
    print("\n[gpu] evolution runs", flush=True)
    ctx = grn.Context(0)
    prob = evolve.scaled_config4(seed=2, n_islands=4, island_size=128, steps=400)
    evolve.make_target(ctx, prob)
    a = evolve.evolve(ctx, prob, False, max_generations=12, threshold=0.0)
    b = evolve.evolve(ctx, prob, False, max_generations=12, threshold=0.0)
    assert a.best == b.best and a.verified and b.verified
    assert all(x >= y for x, y in zip(a.best, a.best[1:]))
    s = evolve.evolve(ctx, prob, True, max_generations=12, threshold=0.0, oversample=4, train_steps=5)
    assert s.verified and all(x >= y for x, y in zip(s.best, s.best[1:]))
    assert s.best[0] == a.best[0]  # the same initial population
    assert all(sc["n_candidates"] in (0, 4 * 512) for sc in s.screen)
    ctx.close()
