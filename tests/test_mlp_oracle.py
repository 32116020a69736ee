"""Pins the FP64 MLP oracle (oracle/mlp_oracle.py) against what the
mathematics fixes, independently of the oracle's own formulas:

* the analytic gradient of the loss against central finite differences
  (every parameter of a small network, FP64);
* GELU' against the numeric derivative of GELU, GELU against x Phi(x);
* Adam's first step against its closed form lr * g / (|g| + eps);
* BF16 rounding against known bit patterns (ties to even);
* the surrogate's encoding of a known model (host code of the library).
"""
import numpy as np
import pytest
from scipy.stats import norm

import grn
import mlp_oracle as mo


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(0)
    width, layers, n = 6, 4, 5
    p = rng.normal(0, 0.6, mo.n_params(width, layers))
    X = rng.normal(0, 1, (n, width))
    t = rng.normal(0, 1, n)
    loss, g = mo.loss_and_grad(p, X, t, width, layers)
    h = 1e-6
    num = np.zeros_like(p)
    for i in range(len(p)):
        e = np.zeros_like(p)
        e[i] = h
        num[i] = (mo.loss_and_grad(p + e, X, t, width, layers)[0] -
                  mo.loss_and_grad(p - e, X, t, width, layers)[0]) / (2 * h)
    assert np.max(np.abs(num - g)) <= 1e-7 * max(1.0, np.max(np.abs(g)))
    assert loss == pytest.approx(np.mean((mo.forward(p, X, width, layers)[0] - t) ** 2), rel=1e-15)


def test_pair_term_matches_finite_differences_and_vanishes_for_equal_rows():
    rng = np.random.default_rng(1)
    width, layers, n, pw = 6, 3, 6, 2.5
    p = rng.normal(0, 0.6, mo.n_params(width, layers))
    X = rng.normal(0, 1, (n, width))
    t = rng.normal(0, 1, n)
    loss, g = mo.loss_and_grad(p, X, t, width, layers, pair_w=pw)
    h = 1e-6
    num = np.zeros_like(p)
    for i in range(len(p)):
        e = np.zeros_like(p)
        e[i] = h
        num[i] = (mo.loss_and_grad(p + e, X, t, width, layers, pair_w=pw)[0] -
                  mo.loss_and_grad(p - e, X, t, width, layers, pair_w=pw)[0]) / (2 * h)
    assert np.max(np.abs(num - g)) <= 1e-7 * max(1.0, np.max(np.abs(g)))
    # the pair term differs from the plain loss: it must change the gradient
    assert np.max(np.abs(g - mo.loss_and_grad(p, X, t, width, layers)[1])) > 1e-3
    # a pair of identical rows predicts no change: its term is the constant
    # pw (t_a - t_b)^2 / n and leaves the gradient as it is without pairs
    X2 = np.repeat(X[:n // 2], 2, axis=0)
    l0, g0 = mo.loss_and_grad(p, X2, t, width, layers)
    l1, g1 = mo.loss_and_grad(p, X2, t, width, layers, pair_w=pw)
    assert np.max(np.abs(g1 - g0)) <= 1e-12 * np.max(np.abs(g0))
    assert l1 - l0 == pytest.approx(pw * np.sum((t[0::2] - t[1::2]) ** 2) / n, rel=1e-12)


def test_gelu_and_derivative():
    z = np.linspace(-6, 6, 1001)
    assert np.allclose(mo.gelu(z), z * norm.cdf(z), rtol=0, atol=1e-14)
    h = 1e-6
    num = (mo.gelu(z + h) - mo.gelu(z - h)) / (2 * h)
    assert np.allclose(mo.gelu_grad(z), num, rtol=0, atol=1e-8)


def test_adam_first_step_closed_form():
    rng = np.random.default_rng(1)
    p = rng.normal(size=100)
    g = rng.normal(size=100)
    p1, m, v = mo.adam_step(p, np.zeros(100), np.zeros(100), g, 1, lr=1e-3, eps=1e-8)
    assert np.allclose(p1, p - 1e-3 * g / (np.abs(g) + 1e-8), rtol=0, atol=1e-15)
    assert np.allclose(m, 0.1 * g) and np.allclose(v, 0.001 * g * g)


def test_bf16_rounding():
    x = np.array([1.0, 1.0 + 2.0 ** -8, 1.0 + 3 * 2.0 ** -8, -2.5, 0.1, 3.0e38])
    bits = mo.bf16_bits(x)
    # 1+2^-8 is a tie between 1.0 (even) and 1+2^-7: rounds to 1.0;
    # 1+3*2^-8 ties between 1+2^-7 (odd) and 1+2^-6 (even): rounds up
    assert list(bits[:4]) == [0x3F80, 0x3F80, 0x3F82, 0xC020]
    assert bits[4] == 0x3DCD  # 0.1 = 0x3DCCCCCD -> 0x3DCD
    assert mo.bf16_round(x)[2] == 1.0 + 2.0 ** -6
    assert np.all(mo.bf16_round(mo.bf16_round(x)) == mo.bf16_round(x))


def test_encoding_of_a_known_model():
    g = np.zeros(4, grn.GENE_DTYPE)
    g["label"] = [0, 1, 2, 11]
    g["lim"] = [0.5, 0.25, 0.75, 1.0]
    g["deg"] = [0.1, 0.2, 0.3, 0.4]
    g["dif"] = [0.0, 0.05, 0.0, 0.01]
    l = np.zeros(4, grn.LINK_DTYPE)
    l[0] = (0, 2, 0.3, 2.0, 0)    # activator OR, slot 0 -> slot 2
    l[1] = (11, 2, 0.6, -4.0, 1)  # repressor, slot 3 -> slot 2
    l[2] = (1, 11, 0.9, 1.0, 1)   # activator AND, slot 1 -> slot 3
    l[3] = (2, 7, 0.5, 1.0, 0)    # to a missing gene: not in use, not encoded
    m = grn.Model(g[[3, 0, 2, 1]], l)  # order of genes in the model is irrelevant
    f = grn.encode(m, 3)
    assert f.shape == (468,)
    want = np.zeros(468, np.float32)
    for s, (lim, deg, dif) in enumerate([(0.5, 0.1, 0.0), (0.25, 0.2, 0.05), (0.75, 0.3, 0.0),
                                         (1.0, 0.4, 0.01)]):
        want[3 * s:3 * s + 3] = (lim, deg, 10 * dif)
    for r, t, typ, K, n in [(0, 2, 1.0, 0.3, 0.5), (3, 2, -1.0, 0.6, 1.0), (1, 3, 0.5, 0.9, 0.25)]:
        b = 36 + 3 * (12 * r + t)
        want[b:b + 3] = (typ, K, n)
    assert np.allclose(f, want, rtol=0, atol=1e-7)
    # a gene outside the in-use closure leaves no trace
    g2 = np.concatenate([g, np.array([(20, 0, 0.9, 0.9, 0.9)], grn.GENE_DTYPE)])
    assert np.array_equal(grn.encode(grn.Model(g2, l), 3), f)
