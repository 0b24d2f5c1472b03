"""The MLP oracle's backward pass against central finite differences of its
own forward (FP64), and its parameter layout against the C ABI's constants."""
from __future__ import annotations

import os
import sys

import numpy as np

sys.path.insert(0, os.path.join(os.path.dirname(__file__), "..", "oracle"))
import mlp_ref as M  # noqa: E402

import alebgk4paper_amd as A  # noqa: E402


def test_layout_matches_abi():
    assert M.TRAIN == A.MLP_TRAINABLE == 4869
    assert M.PARAMS == A.MLP_PARAMS == 4889
    hdr = open(os.path.join(os.path.dirname(__file__), "..", "include", "alebgk_dev.h")).read()
    assert "#define AB_MLP_TRAINABLE 4869" in hdr and "#define AB_MLP_PARAMS 4889" in hdr


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(3)
    p = M.random_params(11)
    n = 17
    x = rng.normal(size=(n, 5))
    t = rng.normal(size=(n, 5))
    w = rng.uniform(0, 1, n); w[3] = 0.0
    L, g = M.loss_grad(p, x, t, w)
    assert abs(L - M.loss(p, x, t, w)) < 1e-15
    idx = np.concatenate([rng.choice(M.TRAIN, 60, replace=False), [M.B1, M.B2, M.B3, M.W3, M.TRAIN - 1]])
    h = 1e-6
    for k in idx:
        pp, pm = p.copy(), p.copy()
        pp[k] += h; pm[k] -= h
        fd = (M.loss(pp, x, t, w) - M.loss(pm, x, t, w)) / (2 * h)
        assert abs(fd - g[k]) <= 1e-7 * max(1.0, abs(g[k])), (k, fd, g[k])


def test_zero_weight_rows_do_not_count():
    rng = np.random.default_rng(5)
    p = M.random_params(2)
    x = rng.normal(size=(8, 5)); t = rng.normal(size=(8, 5))
    w = np.ones(8); w[5:] = 0.0
    L, g = M.loss_grad(p, x, t, w)
    t2 = t.copy(); t2[5:] += 100.0
    L2, g2 = M.loss_grad(p, x, t2, w)
    assert L == L2 and np.array_equal(g, g2)


def test_bf16_rounding_known_values():
    """Round to nearest, ties to even, on the 8-bit significand."""
    one = 1.0
    cases = [(one, one), (one + 2.0 ** -9, one),                 # tie -> even (down)
             (one + 3 * 2.0 ** -9, one + 2.0 ** -7),              # tie -> even (up)
             (one + 2.0 ** -8 + 2.0 ** -20, one + 2.0 ** -7),     # above the tie
             (-3.0, -3.0), (0.0, 0.0), (2.0 ** -130, 2.0 ** -130)]
    for x, want in cases:
        assert M.bf16(np.float32(x)) == want, (x, M.bf16(np.float32(x)), want)


def test_bf16_forward_close_to_fp32_forward():
    """The BF16 inference differs from the FP32 network by the rounding of the
    weights and of the layer-1 activations only: a few 1e-3 of the outputs'
    spread, not more (a misplaced rounding of a bias or of layer 2 would show
    up as a larger error; no rounding at all as zero)."""
    rng = np.random.default_rng(5)
    p = M.random_params(5)
    x = rng.normal(size=(256, M.NIN)).astype(np.float32)
    y = M.forward(p, x)
    yb = M.forward_bf16(p, x)
    d = np.abs(yb - y).max() / np.abs(y).max()
    assert 1e-5 < d < 3e-2, d
