"""Pins the FP64 CPU oracle (oracle/grn_oracle.c) against facts that the
mathematics fixes, independently of the oracle's own code:

* Euler closed forms of pure decay and of Hill-driven production from a
  constant regulator (pins the Hill activation / repression formulas, the
  lim/deg terms and the update rule);
* the reference's fold rule h = min(max_OR H, min_AND H), and h = 0 for a
  gene regulated only by repressors (modelsimulator.cpp:120-138);
* exact conservation and the discrete Neumann eigenmodes of pure diffusion
  (pins the neighbour sums and the no-flux boundary);
* the stop rule / penalty / sentinel / ceilS contract
  (simulator.cpp:54-63, evaluatormorphogens.cpp:48-76, mathalgo.h:55-62);
* the minConc clamp, the rate-zero skip and the diffusion guard of the
  update (modelsimulator.cpp:255,263-275), each with a case a flipped
  comparison would fail;
* invariances: relabelling free genes, genes outside the in-use closure,
  mirror symmetry;
* the experiment-image decode (simstate.cpp:71-90) on pixels with known
  channel bytes.
"""
import math

import numpy as np
import pytest

import grn_oracle as ora

G_DT = ora.GENE_DTYPE
L_DT = ora.LINK_DTYPE


def genes(*rows):
    g = np.zeros(len(rows), G_DT)
    for i, (label, lim, deg, dif) in enumerate(rows):
        g[i] = (label, 0.0, lim, deg, dif)
    return g


def links(*rows):
    l = np.zeros(len(rows), L_DT)
    for i, (r, t, K, n, a) in enumerate(rows):
        l[i] = (r, t, K, n, a)
    return l


def uniform(W, H, *vals):
    return np.stack([np.full((H, W), v, np.float64) for v in vals])[None]


def run(p, init, target, g, l, **kw):
    return ora.evaluate(p, init, target, g, l, want_fields=True, **kw)


# ---- Euler closed forms ----------------------------------------------------

def test_pure_decay_closed_form_and_inputs_evolve():
    # Gene 0 is an input with no regulators: h = 0 so c' = c - dt*deg*c.
    # The reference does not clamp input genes (finding: they evolve).
    W, H, S, dt, deg, c0 = 5, 4, 137, 0.01, 0.2, 0.8
    p = ora.params(W, H, 1, 1, S, dt)
    r, f = run(p, uniform(W, H, c0), uniform(W, H, 0.0), genes((0, 0.3, deg, 0.0), (1, 0.5, 0.1, 0.0)),
               links())
    expect = c0 * (1 - dt * deg) ** S
    assert r["steps"] == S
    np.testing.assert_allclose(f[0], expect, rtol=1e-12)
    assert np.all(f[1] == 0.0)
    assert r["raw_error"] == pytest.approx(0.0, abs=0)  # target plane 0 == field 0


@pytest.mark.parametrize("act", [True, False])
@pytest.mark.parametrize("x,K,n", [(0.3, 0.5, 2.0), (0.7, 0.2, 3.5), (0.5, 0.5, 1.7), (0.9, 0.05, 4.0)])
def test_hill_production_closed_form(act, x, K, n):
    # Input 0 held constant (lim = deg = dif = 0); target gene 1 produced
    # through one Hill link: c_s = c* (1 - (1 - dt*deg)^s), c* = lim*H/deg.
    W, H, S, dt, lim, deg = 3, 3, 250, 0.02, 0.8, 0.35
    p = ora.params(W, H, 1, 1, S, dt)
    l = links((0, 1, K, n if act else -n, 0))
    if not act:
        # a gene with only a repressor gets h = 0; add an activator with
        # H = 1 (K = 0) on the AND side so that h = min(1, H_rep) = H_rep
        l = links((0, 1, K, -n, 1), (0, 1, 0.0, 1.0, 1))
    r, f = run(p, uniform(W, H, x), uniform(W, H, 0.0), genes((0, 0, 0, 0), (1, lim, deg, 0)), l)
    hill = x ** n / (x ** n + K ** n) if act else K ** n / (x ** n + K ** n)
    expect = lim * hill / deg * (1 - (1 - dt * deg) ** S)
    np.testing.assert_allclose(f[1], expect, rtol=1e-12)
    np.testing.assert_allclose(f[0], x, rtol=0)


def test_hill_half_saturation_identity():
    # H_act(K) = H_rep(K) = 1/2 for any n.
    W, H, S, dt, lim, deg = 2, 2, 40, 0.05, 1.0, 0.5
    p = ora.params(W, H, 1, 1, S, dt)
    for n in (1.0, 2.3, 4.0):
        r, f = run(p, uniform(W, H, 0.37), uniform(W, H, 0.0), genes((0, 0, 0, 0), (1, lim, deg, 0)),
                   links((0, 1, 0.37, n, 0)))
        np.testing.assert_allclose(f[1], 0.5 * lim / deg * (1 - (1 - dt * deg) ** S), rtol=1e-12)


def _h_after(link_rows, x0=0.3, x1=0.8):
    # Two constant inputs, target gene 2 with the given links; returns h
    # recovered from the closed form.
    W, H, S, dt, lim, deg = 2, 2, 60, 0.05, 1.0, 0.25
    p = ora.params(W, H, 2, 1, S, dt)
    r, f = run(p, uniform(W, H, x0, x1), uniform(W, H, 0.0),
               genes((0, 0, 0, 0), (1, 0, 0, 0), (2, lim, deg, 0)), links(*link_rows))
    return f[2, 0, 0] * deg / lim / (1 - (1 - dt * deg) ** S)


def hact(x, K, n):
    return x ** n / (x ** n + K ** n)


def hrep(x, K, n):
    return K ** n / (x ** n + K ** n)


def test_fold_or_is_max():
    h = _h_after([(0, 2, 0.4, 2.0, 0), (1, 2, 0.9, 3.0, 0)])
    assert h == pytest.approx(max(hact(0.3, 0.4, 2.0), hact(0.8, 0.9, 3.0)), rel=1e-12)


def test_fold_and_is_min():
    h = _h_after([(0, 2, 0.4, 2.0, 1), (1, 2, 0.9, 3.0, 1)])
    assert h == pytest.approx(min(hact(0.3, 0.4, 2.0), hact(0.8, 0.9, 3.0)), rel=1e-12)


def test_fold_or_then_and():
    # OR activator from 0, AND repressor from 1, AND activator from 1:
    # h = min(max_OR, min_AND) = min(Hact0, Hrep1, Hact1)
    h = _h_after([(1, 2, 0.5, -2.5, 1), (0, 2, 0.2, 1.5, 0), (1, 2, 0.6, 1.2, 1)])
    want = min(hact(0.3, 0.2, 1.5), hrep(0.8, 0.5, 2.5), hact(0.8, 0.6, 1.2))
    assert h == pytest.approx(want, rel=1e-12)


def test_only_repressors_give_zero_production():
    # modelsimulator.cpp:120-138: no activator among a gene's links -> h = 0
    h = _h_after([(0, 2, 0.5, -2.0, 1), (1, 2, 0.7, -1.5, 1)])
    assert h == 0.0


# ---- diffusion ---------------------------------------------------------------

def test_diffusion_conserves_mass():
    W, H, S, dt, D = 9, 7, 300, 0.1, 0.9
    rng = np.random.default_rng(3)
    init = rng.uniform(0.1, 1.0, (1, 1, H, W))
    p = ora.params(W, H, 1, 1, S, dt)
    r, f = run(p, init, np.zeros((1, 1, H, W)), genes((0, 0, 0, D), (1, 0, 0, 0)), links())
    assert f[0].sum() == pytest.approx(init.sum(), rel=1e-13)
    assert f[0].std() < init.std()


@pytest.mark.parametrize("k,l", [(1, 0), (0, 2), (3, 1), (5, 4)])
def test_diffusion_neumann_eigenmode(k, l):
    # cos(pi k (i+1/2)/W) cos(pi l (j+1/2)/H) is an eigenvector of the
    # no-flux 5-point Laplacian with eigenvalue -4(sin^2(pi k/2W) + sin^2(pi l/2H)).
    W, H, S, dt, D, A, B = 12, 10, 200, 0.05, 0.6, 1.0, 0.4
    i = np.arange(W)[None, :]
    j = np.arange(H)[:, None]
    mode = np.cos(np.pi * k * (i + 0.5) / W) * np.cos(np.pi * l * (j + 0.5) / H)
    lam = 1 - dt * D * 4 * (np.sin(np.pi * k / (2 * W)) ** 2 + np.sin(np.pi * l / (2 * H)) ** 2)
    p = ora.params(W, H, 1, 1, S, dt)
    r, f = run(p, (A + B * mode)[None, None], np.zeros((1, 1, H, W)), genes((0, 0, 0, D), (1, 0, 0, 0)),
               links())
    np.testing.assert_allclose(f[0], A + B * lam ** S * mode, rtol=1e-11, atol=1e-13)


# ---- the update's branches: minConc clamp, rate-zero skip, diffusion guard -------

def test_min_conc_clamp_is_strict_and_then_holds_zero():
    # c0 = 1, dt*deg = 1/2: c_s = 2^-s exactly in binary.  With minConc =
    # 1/8 the value 1/8 itself stays (c < minConc is strict) and the next
    # step (1/16 < 1/8) stores 0, which then stays 0 (R = -deg*0 = 0).
    W, H, dt, deg = 3, 2, 0.5, 1.0
    for S, want in ((2, 0.25), (3, 0.125), (4, 0.0), (9, 0.0)):
        p = ora.params(W, H, 1, 1, S, dt, min_conc=0.125)
        r, f = run(p, uniform(W, H, 1.0), uniform(W, H, 0.0), genes((0, 0, deg, 0), (1, 0, 0, 0)), links())
        assert r["steps"] == S
        assert np.all(f[0] == want), (S, f[0])


def test_stop_rule_after_clamp_to_zero():
    # minConcChange = 0: steps 1..4 change by deg*c = 1, 1/2, 1/4, 1/8 > 0;
    # step 4 clamps to 0, so step 5 has R = 0 everywhere, its change is 0
    # and the loop ends (a >= in the stop test would run all S steps).
    p = ora.params(3, 2, 1, 1, 100, 0.5, min_conc=0.125, min_conc_change=0.0, fitness_kind=ora.FITNESS_REF)
    r, f = run(p, uniform(3, 2, 1.0), uniform(3, 2, 0.0), genes((0, 0, 1.0, 0), (1, 0, 0, 0)), links())
    assert r["steps"] == 5 and r["change"] == 0.0
    assert np.all(f[0] == 0.0)


def test_min_conc_clamp_closed_form_crossing():
    # c_s = c0 (1 - dt deg)^s while >= minConc; the first step below it
    # stores 0.  s* = first s with c0 (1 - dt deg)^s < minConc.
    W, H, dt, deg, c0, mc = 2, 2, 0.01, 0.2, 0.8, 0.5
    s_star = math.ceil(math.log(mc / c0) / math.log(1 - dt * deg))
    assert c0 * (1 - dt * deg) ** (s_star - 1) >= mc > c0 * (1 - dt * deg) ** s_star
    for S, want in ((s_star - 1, c0 * (1 - dt * deg) ** (s_star - 1)), (s_star, 0.0), (s_star + 7, 0.0)):
        p = ora.params(W, H, 1, 1, S, dt, min_conc=mc)
        r, f = run(p, uniform(W, H, c0), uniform(W, H, 0.0), genes((0, 0, deg, 0), (1, 0, 0, 0)), links())
        np.testing.assert_allclose(f[0], want, rtol=1e-12, atol=0)


def test_rate_zero_cell_is_not_clamped():
    # A gene with lim = deg = dif = 0 has R = 0 everywhere: the update is
    # skipped (modelsimulator.cpp:265), so a value below minConc survives;
    # without the skip the clamp would store 0.
    p = ora.params(3, 3, 1, 1, 10, 0.1, min_conc=0.1)
    r, f = run(p, uniform(3, 3, 0.05), uniform(3, 3, 0.0), genes((0, 0, 0, 0), (1, 0, 0, 0)), links())
    assert np.all(f[0] == 0.05)


def test_diffusion_guard_brute_force_3x3():
    # One Euler step of a diffusing input gene (lim = deg = 0, dif = 0.5,
    # dt = 0.1, minConc = -10 so nothing is clamped) on a 3x3 field with
    # zero and negative values.  Per cell, from the step-start field: S = sum
    # of the existing 4-neighbours, n = their count; diffusion applies only
    # when S > 0 or (n > 0 and c > 0) (modelsimulator.cpp:255), and then
    # c' = c + dt*dif*(S - n*c) = c + 0.05*(S - n*c); otherwise R = 0 and the
    # cell keeps its value.  Expected values, worked by hand:
    #   (y,x)  c      S      n  guard  c'
    #   (0,0)  0     -0.5    2  off    0       (S < 0, c = 0: c >= 0 would give -0.025)
    #   (0,1)  0.5   -0.5    3  on     0.4     (0.5 + 0.05*(-0.5 - 1.5))
    #   (0,2) -0.5    1.0    2  on    -0.4     (-0.5 + 0.05*(1.0 + 1.0))
    #   (1,0) -1.0    0.0    3  off   -1.0     (S = 0, c < 0: S >= 0 would give -0.85)
    #   (1,1)  0     -0.5    4  off    0
    #   (1,2)  0.5   -0.75   3  on     0.3875  (0.5 + 0.05*(-0.75 - 1.5))
    #   (2,0)  0     -1.5    2  off    0
    #   (2,1) -0.5   -0.25   3  off   -0.5     (no guard at all would give -0.4375)
    #   (2,2) -0.25   0.0    2  off   -0.25
    c = np.array([[0.0, 0.5, -0.5], [-1.0, 0.0, 0.5], [0.0, -0.5, -0.25]])
    want = np.array([[0.0, 0.4, -0.4], [-1.0, 0.0, 0.3875], [0.0, -0.5, -0.25]])
    p = ora.params(3, 3, 1, 1, 1, 0.1, min_conc=-10.0)
    r, f = run(p, c[None, None], np.zeros((1, 1, 3, 3)), genes((0, 0, 0, 0.5), (1, 0, 0, 0)), links())
    np.testing.assert_allclose(f[0], want, rtol=1e-15, atol=1e-16)
    # the guard is read from the step-start field: a second step continues
    # from these values (cell (1,0) now sees S = 0 + 0 + 0 = 0 and c = -1: off)
    p2 = ora.params(3, 3, 1, 1, 2, 0.1, min_conc=-10.0)
    r2, f2 = run(p2, c[None, None], np.zeros((1, 1, 3, 3)), genes((0, 0, 0, 0.5), (1, 0, 0, 0)), links())
    assert f2[0, 1, 0] == -1.0


# ---- contract: stop rule, penalty, sentinel, ceilS -----------------------------

def test_stop_rule_counts_steps():
    # c_s = 0.8*(1-0.002)^s; the step that starts from c has max|R| = 0.2c.
    # The loop stops after the first step whose change is <= 0.1.
    W, H, dt, deg, c0, mcc = 3, 2, 0.01, 0.2, 0.8, 0.1
    p = ora.params(W, H, 1, 1, 100000, dt, min_conc_change=mcc, fitness_kind=ora.FITNESS_REF)
    r, _ = run(p, uniform(W, H, c0), uniform(W, H, 0.0), genes((0, 0, deg, 0), (1, 0, 0, 0)), links())
    s = 1
    while deg * c0 * (1 - dt * deg) ** (s - 1) > mcc:
        s += 1
    assert r["steps"] == s
    assert r["change"] == pytest.approx(deg * c0 * (1 - dt * deg) ** (s - 1), rel=1e-12)
    assert r["raw_error"] == pytest.approx(0.0, abs=0)  # converged: no penalty, target matches


def test_threshold_above_one_runs_no_step():
    # change starts at 1 (simulator.cpp:55), so minConcChange >= 1 stops at once
    p = ora.params(2, 2, 1, 1, 50, 0.01, min_conc_change=1.0, fitness_kind=ora.FITNESS_REF)
    r, f = run(p, uniform(2, 2, 0.5), uniform(2, 2, 0.0), genes((0, 0.3, 0.2, 0), (1, 0, 0, 0)), links())
    assert r["steps"] == 0 and r["change"] == 1.0
    assert r["raw_error"] == 0.0  # state == target and change <= threshold


def test_penalty_with_negative_threshold():
    W, H, S, dt, deg, c0, mcc = 2, 2, 30, 0.01, 0.2, 0.8, -0.5
    p = ora.params(W, H, 1, 1, S, dt, min_conc_change=mcc, fitness_kind=ora.FITNESS_REF)
    r, _ = run(p, uniform(W, H, c0), uniform(W, H, 0.0), genes((0, 0, deg, 0), (1, 0, 0, 0)), links())
    last = deg * c0 * (1 - dt * deg) ** (S - 1)
    assert r["steps"] == S
    assert r["raw_error"] == pytest.approx(last - mcc, rel=1e-12)


def test_log_fitness_and_threshold():
    # REF error = sum log(1 + (|a-b| - thr)) over cells with a positive term / (W*H)
    W, H = 4, 3
    tgt = np.zeros((1, 1, H, W))
    tgt[0, 0, 0, :] = 0.5   # |a-b| = 0.5 on one row
    tgt[0, 0, 1, :2] = 0.1  # |a-b| = 0.1 on two cells
    p = ora.params(W, H, 1, 1, 0, 0.01, min_conc_change=2.0, dist_error_threshold=0.05,
                   fitness_kind=ora.FITNESS_REF)
    r, _ = run(p, uniform(W, H, 0.0), tgt, genes((0, 0, 0, 0), (1, 0, 0, 0)), links())
    want = (W * math.log(1 + 0.45) + 2 * math.log(1 + 0.05)) / (W * H)
    assert r["raw_error"] == pytest.approx(want, rel=1e-14)
    p.dist_error_threshold = 0.6  # above every |a-b|
    r, _ = run(p, uniform(W, H, 0.0), tgt, genes((0, 0, 0, 0), (1, 0, 0, 0)), links())
    assert r["raw_error"] == 0.0


def test_mse_fitness():
    W, H = 4, 3
    rng = np.random.default_rng(0)
    tgt = rng.uniform(0, 1, (1, 1, H, W))
    p = ora.params(W, H, 1, 1, 0, 0.01, fitness_kind=ora.FITNESS_MSE)
    r, _ = run(p, uniform(W, H, 0.0), tgt, genes((0, 0, 0, 0), (1, 0, 0, 0)), links())
    assert r["raw_error"] == pytest.approx(np.mean(tgt ** 2), rel=1e-14)


def test_max_error_sentinel_and_experiment_cut():
    # error accumulates over experiments and the loop stops once it exceeds
    # maxError; the result is then the sentinel 100.
    W, H = 2, 2
    tgt = np.ones((3, 1, H, W))
    p = ora.params(W, H, 1, 1, 0, 0.01, fitness_kind=ora.FITNESS_MSE)
    init = np.zeros((3, 1, H, W))
    r, _ = run(p, init, tgt, genes((0, 0, 0, 0), (1, 0, 0, 0)), links(), max_error=0.5)
    assert r["experiments"] == 2 and r["fitness"] == 100.0
    r, _ = run(p, init, tgt, genes((0, 0, 0, 0), (1, 0, 0, 0)), links(), max_error=1.5)
    assert r["experiments"] == 3 and r["fitness"] == pytest.approx(1.0)


def test_ceil_s_values():
    assert ora.ceil_s(0.0, 5) == 0.0
    assert ora.ceil_s(1.23456, 3) == pytest.approx(1.24, abs=0)
    assert ora.ceil_s(-1.23456, 3) == pytest.approx(-1.23, abs=0)
    assert ora.ceil_s(123456.0, 2) == 130000.0
    kat = 0.048701011019999997
    once = ora.ceil_s(kat, 10)
    assert once == pytest.approx(0.04870101102, rel=1e-15)
    assert ora.ceil_s(once, 10) == once


# ---- invariances -----------------------------------------------------------------

def _random_case(seed, W=6, H=5, S=80):
    rng = np.random.default_rng(seed)
    g = genes((0, rng.uniform(), rng.uniform(0.01, 1), rng.uniform(0, 0.1)),
              (1, rng.uniform(), rng.uniform(0.01, 1), rng.uniform(0, 0.1)),
              (5, rng.uniform(), rng.uniform(0.01, 1), rng.uniform(0, 0.1)))
    l = links((0, 1, 0.4, 2.2, 0), (5, 1, 0.3, -1.5, 1), (1, 5, 0.6, 3.1, 1), (0, 5, 0.2, 1.3, 0))
    init = rng.uniform(0, 1, (1, 1, H, W))
    tgt = rng.uniform(0, 1, (1, 1, H, W))
    return ora.params(W, H, 1, 1, S, 0.05), init, tgt, g, l


def test_relabel_free_gene_is_bitwise_invariant():
    p, init, tgt, g, l = _random_case(1)
    r0, f0 = run(p, init, tgt, g, l)
    g2, l2 = g.copy(), l.copy()
    g2["label"][g2["label"] == 5] = 97
    for key in ("regulator", "regulated"):
        l2[key][l2[key] == 5] = 97
    r1, f1 = run(p, init, tgt, g2, l2)
    assert r0["raw_error"] == r1["raw_error"]
    assert np.array_equal(f0, f1)


def test_genes_outside_closure_change_nothing():
    p, init, tgt, g, l = _random_case(2)
    r0, f0 = run(p, init, tgt, g, l)
    g2 = np.concatenate([g, genes((40, 0.9, 0.1, 0.05))])
    l2 = np.concatenate([l, links((1, 40, 0.5, 2.0, 0), (40, 40, 0.5, 1.0, 0))])  # 40 regulates nothing in use
    assert list(ora.labels_in_use(g2, l2, 2)) == [0, 1, 5]
    r1, f1 = run(p, init, tgt, g2, l2)
    assert r0["raw_error"] == r1["raw_error"]
    assert np.array_equal(f0, f1)


def test_mirror_symmetry():
    p, init, tgt, g, l = _random_case(3)
    r0, f0 = run(p, init, tgt, g, l)
    r1, f1 = run(p, init[..., ::-1].copy(), tgt, g, l)
    assert np.array_equal(f0[..., ::-1], f1)  # left/right sums commute exactly
    r2, f2 = run(p, init[..., ::-1, :].copy(), tgt, g, l)
    np.testing.assert_allclose(f0[:, ::-1, :], f2, rtol=1e-13, atol=1e-15)


def test_kat_model_in_use_closure():
    # the reference comments the known-answer model as a "4-product model"
    # (errorcalculatormultithread.cpp:60-61) with labels 0, 1 structural
    import grn
    m = grn.Model.from_text(open(__file__.replace("test_oracle.py", "golden/kat_model.txt")).read().strip())
    assert list(ora.labels_in_use(m.genes.astype(G_DT), m.links.astype(L_DT), 2)) == [0, 1, 64653, 107652]


def test_batch_matches_single_model_calls():
    # the batch binding packs models into one gene / link array; a packing
    # mistake (e.g. a dropped dtype padding) shows up as a different result
    import workloads
    w = workloads.config0(seed=4, n_models=6)
    w.params.num_sim_steps = 20
    p = ora.from_simparams(w.params)
    tgt = np.random.default_rng(0).uniform(0, 1, (1, 1, 32, 32))
    batch = ora.evaluate_batch(p, w.init, tgt, w.models, n_threads=3)
    single = [ora.evaluate(p, w.init, tgt, m.genes, m.links)[0]["raw_error"] for m in w.models]
    assert np.array_equal(batch["raw_error"], single)


# ---- experiment images -------------------------------------------------------------

def test_decode_rgb_channels_and_y_flip():
    # SimState::loadMorphologyImage (simstate.cpp:71-90): plane k is byte
    # 2 - k of the pixel (red, green, blue) / 255; image scan line j (top
    # first) lands in plane row H - 1 - j.
    W, H = 3, 2
    img = np.array([[[0xFF102030, 0x00FF0000, 0x0000FF00],
                     [0x000000FF, 0x00808080, 0xFFFFFFFF]]], np.uint32)
    out = ora.decode_rgb(img, 3)
    assert out.shape == (1, 3, H, W)
    # top scan line -> row 1, bottom scan line -> row 0
    np.testing.assert_array_equal(out[0, :, 1, 0], [0x10 / 255, 0x20 / 255, 0x30 / 255])
    np.testing.assert_array_equal(out[0, :, 1, 1], [1.0, 0.0, 0.0])
    np.testing.assert_array_equal(out[0, :, 1, 2], [0.0, 1.0, 0.0])
    np.testing.assert_array_equal(out[0, :, 0, 0], [0.0, 0.0, 1.0])
    np.testing.assert_array_equal(out[0, :, 0, 1], [128 / 255] * 3)
    np.testing.assert_array_equal(out[0, :, 0, 2], [1.0, 1.0, 1.0])
    # fewer planes read only the leading channels; the alpha byte is ignored
    np.testing.assert_array_equal(ora.decode_rgb(img, 1)[0, 0], out[0, 0])
    with pytest.raises(ValueError):
        ora.decode_rgb(img, 4)


def test_ranking_check_orders_only_distinct_fitness():
    # the parity tests' ranking comparison: exact and near ties carry no
    # order, every pair the reference tells apart must keep its order
    from parity_util import same_ranking
    want = np.array([0.5, 0.2, 0.2, 0.9, 0.5 * (1 + 1e-7)])
    assert same_ranking(want * (1 + 1e-12), want)
    assert same_ranking(np.array([0.5, 0.2 + 1e-12, 0.2, 0.9, 0.5]), want)   # ties may flip
    assert not same_ranking(np.array([0.5, 0.95, 0.2, 0.9, 0.5]), want)     # 0.2 < 0.9 reversed
