"""GPU parity of the 2-D table transport (AB_PREC_FP64_TABLE with dim 2):
the 2-D Chu step with the MLS transport as per-class coefficient tables
(Solver/CudaBackend.cu:199-300 factorised once per lattice class and dt).

Tolerances (stated per check):
  transport vs the oracle          1e-11 of each particle's largest |g|
                                   (the same elimination, applied to unit
                                   right-hand sides and summed per offset)
  steps vs the FP64 GPU path       1e-10 relative on g, rho, u, T
  mass                             1e-12 relative
"""
from __future__ import annotations

import numpy as np
import pytest

import alebgk4paper_amd as A
import oracle_ctypes as O
from helpers import rel_err, setup

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


def _dt(nx):
    return 0.5 * (1e-6 / (nx - 1)) / 1303.0


@pytest.mark.parametrize("nx,nv,order", [(10, 6, 1), (12, 8, 2), (31, 16, 2)])
def test_table2d_transport_matches_oracle(nx, nv, order):
    _say(f"[table2d transport] cavity2d nx={nx} nv={nv} order={order}")
    case, cl, ocl, neigh, nn = setup("cavity2d", nx, nv, order=order)
    dt = _dt(nx)
    sol = A.Solver(cl, precision=A.PREC_FP64_TABLE)
    sol.transport_only(dt)
    gt = sol.gt()
    want = O.transport(ocl, cl.g, dt, cl.cfg.radius, order)
    scale = np.abs(cl.g).max(axis=1, keepdims=True)
    err = (np.abs(gt - want) / scale).max()
    assert err < 1e-11, err


@pytest.mark.parametrize("nx,nv,order,nsteps", [(12, 8, 2, 5), (16, 8, 1, 5), (31, 16, 2, 20)])
def test_table2d_steps_match_fp64(nx, nv, order, nsteps):
    _say(f"[table2d steps] cavity2d nx={nx} nv={nv} order={order} steps={nsteps}")
    case, cl, ocl, neigh, nn = setup("cavity2d", nx, nv, order=order)
    dt = _dt(nx)
    ref = A.Solver(cl, precision=A.PREC_FP64)
    tab = A.Solver(cl, precision=A.PREC_FP64_TABLE)
    ref.steps(dt, nsteps)
    tab.steps(dt, nsteps)
    assert rel_err(tab.g(), ref.g()) < 1e-10
    mr, mt = ref.moments(), tab.moments()
    for c in (0, 1, 2, 4):
        assert rel_err(mt[:, c], mr[:, c], floor=1e-300) < 1e-10, c
    assert abs(tab.total_mass() - ref.total_mass()) < 1e-12 * ref.total_mass()


def test_table2d_dt_change_rebuilds():
    """A new dt rebuilds the coefficients: two steps at different dt match
    the FP64 path."""
    _say("[table2d dt] cavity2d 12^2 nv8, two dt")
    case, cl, ocl, neigh, nn = setup("cavity2d", 12, 8, order=2)
    ref = A.Solver(cl, precision=A.PREC_FP64)
    tab = A.Solver(cl, precision=A.PREC_FP64_TABLE)
    for dt in (_dt(12), 0.5 * _dt(12)):
        ref.step(dt)
        tab.step(dt)
    assert rel_err(tab.g(), ref.g()) < 1e-10


def test_table2d_refuses_a_cloud_off_the_lattice():
    """An interior particle moved off its lattice site has its own stencil:
    the setup refuses the cloud instead of applying its class's table."""
    _say("[table2d refusal] cavity2d 12^2 nv8, one particle moved")
    case, cl, ocl, neigh, nn = setup("cavity2d", 12, 8, order=2)
    p = 5 * 12 + 6
    cl.x = cl.x.copy()
    cl.x[p] += 1e-3 * (1e-6 / 11)
    with pytest.raises(A.ALEBGKError, match="lattice class"):
        A.Solver(cl, precision=A.PREC_FP64_TABLE)
    # the FP64 path takes the same cloud
    A.Solver(cl, precision=A.PREC_FP64).close()


def test_table2d_mixed_stays_refused():
    _say("[table2d refusal] cavity2d AB_PREC_MIXED")
    case, cl, ocl, neigh, nn = setup("cavity2d", 10, 6, order=2)
    with pytest.raises(A.ALEBGKError):
        A.Solver(cl, precision=A.PREC_MIXED)


@pytest.mark.parametrize("nx", [13, 31])
def test_table2d_run_lengths_bit_identical(nx):
    """The apply's run lengths (4, 8, 12 outputs per item) sum every output
    in the same order: forced or autotuned, gt is bit-identical. nx=13
    leaves ragged runs (9 interior rows per column)."""
    _say(f"[table2d run lengths] cavity2d nx={nx} nv=16")
    case, cl, ocl, neigh, nn = setup("cavity2d", nx, 16, order=2)
    dt = _dt(nx)
    sol = A.Solver(cl, precision=A.PREC_FP64_TABLE)
    sol.transport_only(dt)
    want = sol.gt()
    for shape in (0, 1, 2, -1):
        sol.set_transport_shape(shape)
        sol.transport_only(dt)
        assert np.array_equal(sol.gt(), want), shape
    table, chosen = sol.autotune(dt, reps=2)
    assert [name for _, name, _ in table] == ["R4", "R8", "R12"]
    assert 0 <= chosen < 3
    sol.transport_only(dt)
    assert np.array_equal(sol.gt(), want)
    with pytest.raises(A.ALEBGKError):
        sol.set_transport_shape(3)


def test_table2d_graph_refuses_new_run_length():
    """A graph holds the apply's work items: a new run length after the
    capture makes the replay refuse instead of reading freed items."""
    _say("[table2d graph] cavity2d 12^2 nv8, run length changed after capture")
    case, cl, ocl, neigh, nn = setup("cavity2d", 12, 8, order=2)
    sol = A.Solver(cl, precision=A.PREC_FP64_TABLE)
    g = A.Graph([sol], _dt(12), 1)
    g.launch()
    g.synchronize()
    sol.set_transport_shape(0)
    with pytest.raises(A.ALEBGKError, match="set up again"):
        g.launch()
    g.close()


def test_table2d_graph_rebuilds_coefficients_for_its_dt():
    """An eager step at another dt rebuilds the coefficients in place; the
    next replay builds them for the graph's dt again, so eager and graphed
    sequences agree bit for bit."""
    _say("[table2d graph] cavity2d 12^2 nv8, eager step at another dt between replays")
    case, cl, ocl, neigh, nn = setup("cavity2d", 12, 8, order=2)
    dt = _dt(12)
    a = A.Solver(cl, precision=A.PREC_FP64_TABLE)
    b = A.Solver(cl, precision=A.PREC_FP64_TABLE)
    g = A.Graph([a], dt, 1)
    a.step(0.5 * dt)
    g.launch()
    g.synchronize()
    b.step(0.5 * dt)
    b.step(dt)
    assert np.array_equal(a.g(), b.g())
    g.close()


def test_table2d_step_with_positivity_matches_oracle():
    """Negative values planted in g1 of a few particles make the positivity
    limiter fire after the transport. The table step's fused commit, wall
    and collision (with the first mass rescale folded into the last) flags
    the same particles as the oracle, and g, the moments and the step's
    masses agree with the oracle and with the FP64 path."""
    _say("[table2d positivity] cavity2d 16^2 nv8, negatives planted in 24 rows")
    case, cl, ocl, neigh, nn = setup("cavity2d", 16, 8, order=2)
    nodes = 64
    g = cl.g.copy()
    rng = np.random.default_rng(11)
    rows = rng.choice(cl.N, 24, replace=False)
    g[np.ix_(rows, np.arange(0, nodes, 5))] *= -0.3
    cl.g = g
    dt = _dt(16)
    tab = A.Solver(cl, precision=A.PREC_FP64_TABLE)
    ref = A.Solver(cl, precision=A.PREC_FP64)
    st = O.State(ocl, cl.g, cl.mom)
    fired = 0
    for k in range(3):
        stats = tab.step(dt, stats=True)
        rstats = ref.step(dt, stats=True)
        m_old = O.step(ocl, st, dt, case.tau, case.R, cl.cfg.radius, 2)
        assert abs(stats["mass_old"] - m_old) <= 1e-12 * m_old, k
        for key in ("mass_after_transport", "mass_after_collision"):
            assert abs(stats[key] - rstats[key]) <= 1e-12 * rstats[key], (k, key)
        fl = st.limflag.astype(np.int64)
        assert stats["n_positivity"] == int(np.count_nonzero(fl & 1)), k
        fired += stats["n_positivity"]
    assert fired > 0
    mom, lim = tab.moments(with_limflag=True)
    assert rel_err(tab.g(), st.g) < 1e-10
    for c in (0, 1, 2, 4):
        assert rel_err(mom[:, c], st.mom[:, c], floor=1e-300) < 1e-10, c
    assert np.array_equal(lim, st.limflag)


def test_table2d_step_without_mass_fix_matches_oracle():
    """With the end-of-step rescale switched off, the table step still applies
    the first rescale (massOld / massNew) that its fused commit and collision
    leave out: g, the moments (the wall densities too) and the flags agree
    with the oracle's step without the fix."""
    _say("[table2d no mass fix] cavity2d 12^2 nv8, 3 steps")
    case, cl, ocl, neigh, nn = setup("cavity2d", 12, 8, order=2)
    dt = _dt(12)
    tab = A.Solver(cl, precision=A.PREC_FP64_TABLE, no_mass_fix=True)
    st = O.State(ocl, cl.g, cl.mom)
    for k in range(3):
        tab.step(dt)
        O.step(ocl, st, dt, case.tau, case.R, cl.cfg.radius, 2, nofix=1)
    mom, lim = tab.moments(with_limflag=True)
    assert rel_err(tab.g(), st.g) < 1e-10
    for c in (0, 1, 2, 4):
        assert rel_err(mom[:, c], st.mom[:, c], floor=1e-300) < 1e-10, c
    assert np.array_equal(lim, st.limflag)
