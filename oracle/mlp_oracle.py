"""FP64 CPU restatement of the surrogate MLP -- TEST INFRASTRUCTURE ONLY.

Only tests/ may import this module; the library never calls it.

The surrogate has no counterpart in the reference (which simulates every
child), so there is nothing of the reference's to follow here: this is the
textbook network the device implements --
    z_l = h_{l-1} W_l^T + b_l,  h_l = GELU(z_l) = z_l Phi(z_l)   (l = 1..L)
    y   = h_L . w + b,          loss = mean over rows of (y - t)^2
with its analytic backward pass and Adam (Kingma & Ba, 2015, Algorithm 1).
Parity pinned by tests/test_mlp_oracle.py: central finite differences of
the loss against the analytic gradient (FP64, small widths), the GELU
derivative against its numeric derivative, Adam's first step against its
closed form, and round-to-nearest-even BF16 against known bit patterns.

With bf16=True the functions round exactly the quantities the device keeps
in BF16 -- inputs, weights (both orientations), hidden activations h_l and
the backward signals dz_l -- and keep everything else in FP64, so that a
device result can be compared against it with a tolerance set by FP32
accumulation alone.  Flat parameter layout (as on the device): W_l^T
(out x in) for l = 1..L, then b_1..b_L, then w, then b.
"""
from __future__ import annotations

import numpy as np
from scipy.special import erf

SQRT_HALF = np.sqrt(0.5)
INV_SQRT_2PI = 1.0 / np.sqrt(2.0 * np.pi)


def bf16_round(x) -> np.ndarray:
    """Round to the nearest BF16 value (ties to even), returned as float64."""
    f = np.asarray(x, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def bf16_bits(x) -> np.ndarray:
    """BF16 bit patterns (uint16) of x, rounded to nearest even."""
    f = np.asarray(x, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    return ((u + 0x7FFF + ((u >> 16) & 1)) >> 16).astype(np.uint16)


def gelu(z):
    return 0.5 * z * (1.0 + erf(z * SQRT_HALF))


def gelu_grad(z):
    return 0.5 * (1.0 + erf(z * SQRT_HALF)) + z * INV_SQRT_2PI * np.exp(-0.5 * z * z)


def n_params(width: int, layers: int) -> int:
    return layers * width * width + layers * width + width + 1


def unpack(p, width: int, layers: int):
    p = np.asarray(p, dtype=np.float64)
    mat = width * width
    Wt = [p[l * mat:(l + 1) * mat].reshape(width, width) for l in range(layers)]
    ob = layers * mat
    b = [p[ob + l * width: ob + (l + 1) * width] for l in range(layers)]
    ow = ob + layers * width
    return Wt, b, p[ow:ow + width], p[ow + width]


def forward(p, X, width: int, layers: int, bf16: bool = False):
    """(y, cache) for rows X (n x width)."""
    Wt, b, w, b0 = unpack(p, width, layers)
    q = bf16_round if bf16 else (lambda a: np.asarray(a, dtype=np.float64))
    h = q(X)
    Z, H = [], [h]
    for l in range(layers):
        z = h @ q(Wt[l]).T + b[l]
        h = q(gelu(z))
        Z.append(z)
        H.append(h)
    y = h @ q(w) + b0
    return y, (Z, H)


def loss_and_grad(p, X, t, width: int, layers: int, bf16: bool = False, pair_w: float = 0.0):
    """Mean squared error over the rows and its gradient (flat layout).
    pair_w > 0: rows 2i and 2i + 1 are a pair and the loss adds
    pair_w / n * sum over pairs of ((y_2i - y_2i+1) - (t_2i - t_2i+1))^2."""
    Wt, b, w, b0 = unpack(p, width, layers)
    q = bf16_round if bf16 else (lambda a: np.asarray(a, dtype=np.float64))
    y, (Z, H) = forward(p, X, width, layers, bf16)
    n = len(t)
    t = np.asarray(t, dtype=np.float64)
    e = y - t
    loss = float(np.mean(e * e))
    dy = 2.0 * e / n
    if pair_w > 0:
        assert n % 2 == 0
        d = (y[0::2] - y[1::2]) - (t[0::2] - t[1::2])  # per pair, seen from its first row
        loss += pair_w * float(np.sum(d * d)) / n
        dy[0::2] += 2.0 * pair_w * d / n
        dy[1::2] -= 2.0 * pair_w * d / n
    g = np.zeros(n_params(width, layers))
    mat = width * width
    ob, ow = layers * mat, layers * mat + layers * width
    g[ow:ow + width] = H[layers].T @ dy
    g[ow + width] = dy.sum()
    dz = q(dy[:, None] * q(w)[None, :] * gelu_grad(Z[layers - 1]))
    for l in range(layers - 1, -1, -1):
        g[l * mat:(l + 1) * mat] = (dz.T @ H[l]).ravel()
        g[ob + l * width: ob + (l + 1) * width] = dz.sum(axis=0)
        if l > 0:
            dz = q((dz @ q(Wt[l])) * gelu_grad(Z[l - 1]))
    return loss, g


def adam_step(p, m, v, g, step: int, lr=1e-3, beta1=0.9, beta2=0.999, eps=1e-8):
    """One Adam update (Algorithm 1 of Kingma & Ba); returns new (p, m, v)."""
    m = beta1 * m + (1 - beta1) * g
    v = beta2 * v + (1 - beta2) * g * g
    mh = m / (1 - beta1 ** step)
    vh = v / (1 - beta2 ** step)
    return p - lr * mh / (np.sqrt(vh) + eps), m, v
