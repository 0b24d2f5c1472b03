"""GPU: the discrete-Maxwellian collision against the oracle, the MLP's HIP
forward and backward kernels against the numpy oracle, and what the learned
warm start gains.

Tolerances:
  FP64 step with the discrete collision vs oracle     1e-10 relative
  mixed vs FP64 (both discrete)                       1e-5 on rho, u, T
  MLP forward / loss / gradient (FP32 kernels)        1e-4 relative to the largest entry
"""
from __future__ import annotations

import os
import sys

import numpy as np
import pytest

import alebgk4paper_amd as A
import oracle_ctypes as O
from helpers import moment_errors, rel_err, setup

sys.path.insert(0, os.path.join(os.path.dirname(__file__), "..", "oracle"))
import mlp_ref as M  # noqa: E402

pytestmark = pytest.mark.gpu
DISC = A.COLLISION_DISCRETE


def _say(msg):
    print(msg, flush=True)


@pytest.mark.parametrize("nx,nv,order,prec", [(8, 6, 2, A.PREC_FP64), (7, 8, 1, A.PREC_FP64),
                                              (8, 6, 2, A.PREC_FP64_TABLE),
                                              (9, 7, 1, A.PREC_FP64_TABLE)])
def test_dm_fp64_step_matches_oracle(nx, nv, order, prec):
    """The direct FP64 path sums the Newton Jacobian node by node; the FP64
    table path from the separable 1-D sums and the collision from per-axis
    factor tables. Both against the oracle's node-by-node FP64."""
    _say(f"[dm] fp64 step vs oracle cavity3d nx={nx} nv={nv} order={order} precision={prec}")
    case, cl, ocl, neigh, nn = setup("cavity3d", nx, nv, order=order)
    dt = 0.5 * (1e-6 / (nx - 1)) / 1303.0
    sol = A.Solver(cl, precision=prec, collision=DISC)
    st = O.State(ocl, cl.g, cl.mom)
    for _ in range(3):
        stats = sol.step(dt, stats=True)
        O.step_dm(ocl, st, dt, case.tau, case.R, cl.cfg.radius, order)
    assert stats["dm_unconverged"] == 0 and stats["dm_res_max"] <= 1e-12
    assert stats["dm_particles"] == int((cl.boundary == 0).sum())
    a, passes, res = sol.dm_download()
    inner = cl.boundary == 0
    assert np.all(passes[~inner] == 0) and passes[inner].min() >= 1
    assert np.all(np.isfinite(a[inner]))
    assert rel_err(sol.g(), st.g) < 1e-10
    mom = sol.moments()
    for c in range(5):
        assert rel_err(mom[:, c], st.mom[:, c], floor=1e-300) < 1e-10, c


@pytest.mark.parametrize("nx,nv,nsteps", [(12, 8, 10)])
def test_dm_mixed_within_1e5_of_fp64(nx, nv, nsteps):
    _say(f"[dm] mixed vs fp64 (discrete) cavity3d nx={nx} nv={nv} steps={nsteps}")
    case, cl, ocl, neigh, nn = setup("cavity3d", nx, nv, order=2)
    ref = A.Solver(cl, precision=A.PREC_FP64, collision=DISC)
    mix = A.Solver(cl, precision=A.PREC_MIXED, collision=DISC)
    for _ in range(nsteps):
        ref.step(case.dt)
        st = mix.step(case.dt, stats=True)
    assert st["dm_unconverged"] == 0
    e = moment_errors(mix.moments(), ref.moments(), cl.boundary == 0)
    _say(f"  errors {e}")
    assert max(e.values()) < 1e-5, e


def _mlp_solver():
    case, cl, ocl, neigh, nn = setup("cavity3d", 6, 5, order=1)
    return A.Solver(cl, precision=A.PREC_FP64, collision=DISC)


def test_mlp_forward_and_gradient_match_oracle():
    _say("[mlp] forward / loss / gradient vs numpy oracle")
    sol = _mlp_solver()
    rng = np.random.default_rng(9)
    p = M.random_params(21).astype(np.float32)
    sol.mlp_set(p)
    assert np.array_equal(sol.mlp_get(), p)
    n = 300
    x = rng.normal(size=(n, 5)).astype(np.float32)
    t = rng.normal(size=(n, 5)).astype(np.float32)
    w = (rng.uniform(size=n) > 0.2).astype(np.float32)
    y = sol.mlp_forward(x)
    y_ref = M.forward(p.astype(np.float64), x.astype(np.float64))
    assert np.abs(y - y_ref).max() <= 1e-4 * np.abs(y_ref).max()
    L, g = sol.mlp_loss_grad(x, t, w)
    L_ref, g_ref = M.loss_grad(p.astype(np.float64), x.astype(np.float64),
                               t.astype(np.float64), w.astype(np.float64))
    assert abs(L - L_ref) <= 1e-4 * abs(L_ref)
    assert np.abs(g - g_ref).max() <= 1e-4 * np.abs(g_ref).max()
    # per block of the parameter vector, so a transposed layer cannot hide
    for a, b in ((M.W1, M.B1), (M.B1, M.W2), (M.W2, M.B2), (M.B2, M.W3), (M.W3, M.B3), (M.B3, M.TRAIN)):
        assert np.abs(g[a:b] - g_ref[a:b]).max() <= 1e-4 * np.abs(g_ref[a:b]).max() + 1e-12, (a, b)


def test_mlp_mfma_inference_matches_bf16_oracle():
    """The warm start's inference (BF16 weights on MFMA, BF16 layer-1
    activations, FP32 accumulation, one wave per 64 rows) against the oracle
    emulating the same roundings. A layout mistake in the MFMA operands gives
    errors of the outputs' own size; what remains here is FP32 accumulation
    order and the rare layer-1 activation that rounds to the other BF16
    neighbour. 300 rows: four full waves and a ragged fifth."""
    _say("[mlp] MFMA inference vs BF16-emulating oracle")
    sol = _mlp_solver()
    rng = np.random.default_rng(11)
    p = M.random_params(23).astype(np.float32)
    sol.mlp_set(p)
    x = rng.normal(size=(300, 5)).astype(np.float32)
    y = sol.mlp_infer(x)
    y_ref = M.forward_bf16(p, x)
    scale = np.abs(y_ref).max()
    err = np.abs(y - y_ref) / scale
    _say(f"  max {err.max():.2e} median {np.median(err):.2e}")
    assert np.median(err) <= 1e-5
    assert err.max() <= 5e-3
    # and it is the BF16 network, not the FP32 one
    assert np.abs(sol.mlp_forward(x) - y).max() > 1e-5 * scale


def test_learned_warm_start_gain():
    """Train the MLP on rows the plain discrete solve generates, then run the
    same steps with and without the warm start: every solve still converges
    to the same tolerance (the moments agree to round-off of the tolerance),
    and the warm start needs fewer grid passes."""
    _say("[mlp] warm-start gain, mixed cavity3d 16^3 nv8")
    case = A.configure("cavity3d", nx=16, nv=8, mls_order=2)
    cl = A.build_lattice(case, precision=A.PREC_MIXED, mls_order=2)
    plain = A.Solver(cl, precision=A.PREC_MIXED, collision=DISC)
    learn = A.Solver(cl, precision=A.PREC_MIXED, collision=DISC)
    for s in (plain, learn):
        s.steps(case.dt, 5)
    learn.dm_record(8)
    learn.steps(case.dt, 8)
    plain.steps(case.dt, 8)
    rows, valid = learn.dm_rows()
    assert rows == 8 * cl.N and valid == 8 * int((cl.boundary == 0).sum())
    learn.mlp_init(seed=1)
    l0, l1 = learn.mlp_train(400, batch=2048, lr=2e-3, seed=5)
    _say(f"  loss {l0:.3e} -> {l1:.3e}")
    assert l1 < 0.1 * l0
    learn.mlp_enable(True)
    tot_p = tot_l = 0
    for _ in range(5):
        sp = plain.step(case.dt, stats=True)
        sl = learn.step(case.dt, stats=True)
        assert sp["dm_unconverged"] == 0 and sl["dm_unconverged"] == 0
        tot_p += sp["dm_passes"]; tot_l += sl["dm_passes"]
    n = 5 * sp["dm_particles"]
    _say(f"  grid passes per solve: plain {tot_p / n:.3f}, warm start {tot_l / n:.3f}")
    assert tot_l < tot_p
    # both solves stop at a relative residual of 1e-12 (momentum in units of
    # the grid speed, 1303 m/s), so the flows agree to ~1e-9 m/s
    e = moment_errors(learn.moments(), plain.moments(), cl.boundary == 0)
    assert max(e.values()) < 1e-7, e


def test_headline_path_500_steps_within_1e5():
    """The measured configuration end to end: cavity3d 40^3, Nv = 10,
    quadratic MLS, lid 1 m/s, 500 steps of the mixed step with the discrete
    collision and the learned warm start (trained on the run's own first 5
    steps), against the FP64 table step with the same collision (pinned to
    the oracle above). rho, u and T agree to 1e-5 relative (|du| relative
    to max |u|), every solve converges, and mass is restored."""
    _say("[headline] mixed+DM+MLP vs fp64-table+DM, cavity3d 40^3 nv10, 500 steps")
    case = A.configure("cavity3d", nx=40, nv=10, mls_order=2)
    cl = A.build_lattice(case, precision=A.PREC_MIXED, mls_order=2)
    dt = case.dt
    mix = A.Solver(cl, precision=A.PREC_MIXED, collision=DISC)
    ref = A.Solver(cl, precision=A.PREC_FP64_TABLE, collision=DISC)
    m0 = ref.total_mass()
    mix.dm_record(5)
    mix.steps(dt, 5)
    ref.steps(dt, 5)
    mix.mlp_init(seed=1)
    l0, l1 = mix.mlp_train(600, batch=4096, lr=2e-3, seed=5)
    _say(f"  MLP loss {l0:.3e} -> {l1:.3e}")
    mix.mlp_enable(True)
    for k in range(5, 500, 99):
        n = min(99, 500 - k)
        mix.steps(dt, n)
        ref.steps(dt, n)
        mix.synchronize()
        ref.synchronize()
        _say(f"  {k + n} steps")
    st = mix.step(dt, stats=True)
    ref.step(dt)
    assert st["dm_unconverged"] == 0
    inner = cl.boundary == 0
    e = moment_errors(mix.moments(), ref.moments(), inner)
    _say(f"  errors after 501 steps {e}")
    assert max(e.values()) < 1e-5, e
    assert abs(mix.total_mass() - m0) < 1e-11 * m0
    assert abs(ref.total_mass() - m0) < 1e-11 * m0
