"""GPU: several contexts' steps captured into one HIP graph give exactly the
eager steps' results (bit-identical), replay after replay, with eager work
and host synchronisation between the replays."""
from __future__ import annotations

import numpy as np
import pytest

import alebgk4paper_amd as A
from helpers import setup

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


def _instances():
    out = []
    for name, nx, nv, prec, order in [("cavity3d", 10, 6, A.PREC_MIXED, 2),
                                      ("cavity3d", 8, 5, A.PREC_FP64, 2),
                                      ("cavity2d", 12, 8, A.PREC_FP64, 2)]:
        case, cl, ocl, neigh, nn = setup(name, nx, nv, order=order, precision=prec)
        out.append((case, cl, prec))
    return out


def test_graph_replays_match_eager_steps():
    _say("[graph] 3 instances (mixed 3-D, FP64 3-D, FP64 2-D), 3 steps per replay")
    inst = _instances()
    dt = 1e-11
    eager = [A.Solver(cl, precision=p) for _, cl, p in inst]
    cap = [A.Solver(cl, precision=p) for _, cl, p in inst]
    graph = A.Graph(cap, dt, 3)
    for rep in range(3):
        graph.launch()
        graph.synchronize()
        for s in eager:
            s.steps(dt, 3)
            s.synchronize()
        for k, (s_e, s_g) in enumerate(zip(eager, cap)):
            assert np.array_equal(s_e.g(), s_g.g()), (rep, k)
            assert np.array_equal(s_e.moments(), s_g.moments()), (rep, k)
    graph.close()


def test_graph_refuses_profiling_context():
    _say("[graph] refuses a profiling context")
    case, cl, prec = _instances()[1]
    s = A.Solver(cl, precision=prec)
    s.profile(True)
    with pytest.raises(A.ALEBGKError):
        A.Graph([s], 1e-11, 1)
    s.profile(False)
    g = A.Graph([s], 1e-11, 1)
    g.launch(); g.synchronize(); g.close()


def test_autotune_keeps_results_and_picks_a_shape():
    """Single-column transport shapes give bit-identical steps (the same
    products in the same order per output); block shapes visit an output's
    stencil columns in another order, so they agree to FP32 rounding. The
    autotuner times every shape that fits and keeps one."""
    _say("[autotune] cavity3d 12^3 nv8 mixed")
    case, cl, ocl, neigh, nn = setup("cavity3d", 12, 8, order=2, precision=A.PREC_MIXED)
    dt = 1e-11
    base = A.Solver(cl, precision=A.PREC_MIXED)
    table, chosen = base.autotune(dt, reps=2)
    _say(f"  {table} -> {chosen}")
    assert chosen in [i for i, _, _ in table] and len(table) >= 8
    assert all(ms > 0 for _, _, ms in table)
    base.set_transport_shape(0)
    base.steps(dt, 3)
    ref = base.g()
    for i, name, _ in table:
        s = A.Solver(cl, precision=A.PREC_MIXED)
        s.set_transport_shape(i)
        s.steps(dt, 3)
        g = s.g()
        if name.startswith("B"):
            assert np.abs(g - ref).max() <= 1e-6 * np.abs(ref).max(), name
        else:
            assert np.array_equal(g, ref), name
        s.close()
