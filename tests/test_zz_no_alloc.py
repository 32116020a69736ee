"""grn_evaluate_batch_noalloc, the entry point that allocates nothing, called
through the C ABI: its results equal grn_evaluate_batch's (on-chip path) on
two calls in a row, the second with fewer models (the shared scratch array
keeps the first call's contents), and it refuses what does not fit."""
import ctypes

import numpy as np
import pytest

import grn
import workloads

pytestmark = pytest.mark.gpu


def datasets(ctx, rng):
    ids = []
    for W, H in ((32, 32), (24, 16), (48, 48)):
        p = grn.SimParams(W, H, 2, 1, 150, 0.05, fitness_kind=grn.FITNESS_MSE)
        init = np.stack([workloads.gradient_x(W, H), workloads.gradient_y(W, H)])[None]
        ids.append(ctx.add_dataset(p, init, rng.uniform(0, 1, (1, 1, H, W))))
    return ids


def call_noalloc(lib, ctx, models, ds, me):
    n = len(models)
    cms = (grn._Model * max(n, 1))(*[m._c() for m in models])
    d = np.ascontiguousarray(ds, np.int32)
    e = np.ascontiguousarray(me, np.float64)
    out = np.zeros(max(n, 1), grn.RESULT_DTYPE)
    rc = lib.grn_evaluate_batch_noalloc(ctx._h, n, d.ctypes.data, ctypes.addressof(cms), e.ctypes.data, None,
                                        out.ctypes.data)
    return rc, out[:n]


def test_noalloc_equals_evaluate_batch_on_two_calls():
    print("\n[gpu] no-alloc entry point", flush=True)
    lib = grn.load_library()
    ctx = grn.Context(0)
    rng = np.random.default_rng(12)
    ids = datasets(ctx, rng)
    models = workloads._population(rng, 240, (3, 9), (1, 24), 3)
    ds = np.array([ids[i % 3] for i in range(240)], np.int32)
    want = ctx.evaluate_batch(models, datasets=ds, path=grn.PATH_LDS)
    cut = np.where(np.arange(240) % 5 == 0, 0.5 * want["raw_error"], np.inf)  # some hit the maxError cut
    want = ctx.evaluate_batch(models, datasets=ds, max_errors=cut, path=grn.PATH_LDS)
    assert np.any(want["fitness"] == 100) and np.any(want["fitness"] != 100)
    print("[gpu] first call, 240 models", flush=True)
    rc, got = call_noalloc(lib, ctx, models, ds, cut)
    assert rc == 0, lib.grn_last_error()
    assert np.array_equal(got, want)
    # a second, smaller call: the array still holds the first call's data
    sub = np.arange(7, 240, 4)[:50]
    print("[gpu] second call, 50 models", flush=True)
    rc, got2 = call_noalloc(lib, ctx, [models[i] for i in sub], ds[sub], cut[sub])
    assert rc == 0, lib.grn_last_error()
    assert np.array_equal(got2, want[sub])
    # the wrapper's method takes the same route
    assert np.array_equal(ctx.evaluate_batch_noalloc([models[i] for i in sub], datasets=ds[sub],
                                                     max_errors=cut[sub]), want[sub])
    # beyond the capacity: refused, nothing runs
    rc, _ = call_noalloc(lib, ctx, models + models[:20], np.concatenate([ds, ds[:20]]), np.full(260, np.inf))
    assert rc == -3 and b"at most 256" in lib.grn_last_error()
    # a grid that needs the halo-tiled kernel: refused
    p = grn.SimParams(128, 128, 2, 1, 10, 0.05, fitness_kind=grn.FITNESS_MSE)
    big = ctx.add_dataset(p, np.zeros((1, 2, 128, 128)), np.zeros((1, 1, 128, 128)))
    m12 = workloads._population(rng, 1, (12, 12), (20, 30), 3)
    rc, _ = call_noalloc(lib, ctx, m12, [big], [np.inf])
    assert rc == -3
    ctx.close()
