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
