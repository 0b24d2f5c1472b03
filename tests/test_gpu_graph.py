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


@pytest.mark.parametrize("nx,nv", [(12, 8), (16, 6)])
def test_autotune_keeps_results_and_picks_a_shape(nx, nv):
    """Single-column transport shapes give bit-identical steps (the same
    products in the same order per output); block shapes visit an output's
    stencil columns in another order, so they agree to FP32 rounding. The
    autotuner times every shape that fits and keeps one. At 16^3 the slab
    shapes' last z block of the edge blocks is a partial one."""
    _say(f"[autotune] cavity3d {nx}^3 nv{nv} mixed")
    case, cl, ocl, neigh, nn = setup("cavity3d", nx, nv, order=2, precision=A.PREC_MIXED)
    dt = 1e-11
    base = A.Solver(cl, precision=A.PREC_MIXED)
    table, chosen = base.autotune(dt, reps=2)
    _say(f"  {table} -> {chosen}")
    assert chosen in [i for i, _, _ in table] and len(table) >= 4
    assert all(ms > 0 for _, _, ms in table)
    base.set_transport_shape(0)
    base.steps(dt, 3)
    ref = base.g()
    blocks = {}
    for i, name, _ in table:
        s = A.Solver(cl, precision=A.PREC_MIXED)
        s.set_transport_shape(i)
        s.steps(dt, 3)
        g = s.g()
        if name.startswith("B"):
            assert np.abs(g - ref).max() <= 1e-6 * np.abs(ref).max(), name
            blocks[name] = g
        else:
            assert np.array_equal(g, ref), name
        s.close()
    # the slab order does the whole-column order's products in its order
    if "B2x2/K12" in blocks and "B2x2/K12/S" in blocks:
        assert np.array_equal(blocks["B2x2/K12"], blocks["B2x2/K12/S"])


def test_eager_step_after_capture_matches_uncaptured():
    """Capturing records work without executing it. A context captured
    straight after its upload and then stepped eagerly (before any replay)
    steps exactly like one that was never captured: the start-of-step mass
    sums are computed before the capture, not marked current by it."""
    _say("[graph] eager step after a capture, FP64 3-D and MOOD")
    case, cl, prec = _instances()[1]
    dt = 1e-11
    for limiter in (A.LIMITER_GPU, A.LIMITER_MOOD):
        plain = A.Solver(cl, precision=prec, limiter=limiter)
        cap = A.Solver(cl, precision=prec, limiter=limiter)
        graph = A.Graph([cap], dt, 2)
        plain.steps(dt, 2)
        cap.steps(dt, 2)
        assert np.array_equal(plain.g(), cap.g()), limiter
        # a replay continues from the eager state exactly as eager steps do
        graph.launch(); graph.synchronize()
        plain.steps(dt, 2)
        assert np.array_equal(plain.g(), cap.g()), limiter
        graph.close()


def test_graph_refuses_replay_after_cloud_changes():
    """A graph holds its clouds' device addresses: a re-upload or a cloud
    management pass (which re-lays the cloud out) invalidates it, and the
    replay is refused instead of running on freed memory."""
    _say("[graph] refused after upload / cloud management")
    case, cl, prec = _instances()[1]
    s = A.Solver(cl, precision=prec)
    g = A.Graph([s], 1e-11, 1)
    g.launch(); g.synchronize()
    s.upload(cl)
    with pytest.raises(A.ALEBGKError):
        g.launch()
    g.close()
    g = A.Graph([s], 1e-11, 1)
    g.launch(); g.synchronize()
    dx = 1e-6 / (case.nx - 1)
    dom = (case.dom[0], case.dom[1], case.dom[3], case.dom[2], case.dom[4], case.dom[5])
    # a sphere in the middle of the box overruns the particles inside it
    src, counts = s.cloud_manage([A.body(A.BODY_SPHERE, (5e-7, 5e-7, 5e-7), 2.2 * dx)],
                                 np.full(s.N, -1, np.int32), dx, 0.0, dom)
    assert counts[0] > 0
    with pytest.raises(A.ALEBGKError):
        g.launch()
    g.close()
    # a graph made afterwards runs
    g = A.Graph([s], 1e-11, 1)
    g.launch(); g.synchronize(); g.close()


def test_batch_scatter_refuses_changed_member():
    _say("[batch] scatter refused after a member re-upload")
    case, cl, prec = _instances()[2]
    a = A.Solver(cl, precision=prec)
    b = A.Solver(cl, precision=prec)
    bt = A.Batch([a, b])
    bt.step(1e-13)
    bt.scatter()
    a.upload(cl)
    with pytest.raises(A.ALEBGKError):
        bt.scatter()
    bt.close()


def test_graph_launch_is_ordered_with_the_contexts_streams():
    """A replay runs after the work already enqueued on its contexts'
    streams and before what is enqueued on them afterwards: eager steps,
    replays and a download interleaved with no synchronisation between them
    give exactly the eager steps' results."""
    _say("[graph] ordering of replays against eager work, cavity3d 16^3 nv6 FP64")
    case, cl, ocl, neigh, nn = setup("cavity3d", 16, 6, order=2)
    dt = case.dt
    plain = A.Solver(cl, precision=A.PREC_FP64)
    mixed = A.Solver(cl, precision=A.PREC_FP64)
    graph = A.Graph([mixed], dt, 2)
    plain.steps(dt, 6)
    mixed.steps(dt, 1)          # eager, then a replay, then eager, then a replay
    graph.launch()
    mixed.steps(dt, 1)
    graph.launch()
    g = mixed.g()               # downloads on the context's stream: after the replay
    assert np.array_equal(g, plain.g())
    graph.close()


def test_graph_rebuilds_table_coefficients_for_its_dt():
    """The captured table step applies coefficients it does not build: after
    an eager step at another dt, the replay builds them for the graph's dt
    first, so eager and graphed sequences agree bit for bit."""
    _say("[graph] mixed 3-D 10^3 nv6, eager step at another dt before a replay")
    case, cl, prec = _instances()[0]
    dt = 1e-11
    a = A.Solver(cl, precision=prec)
    b = A.Solver(cl, precision=prec)
    g = A.Graph([a], dt, 1)
    a.step(0.5 * dt)
    g.launch()
    g.synchronize()
    b.step(0.5 * dt)
    b.step(dt)
    assert np.array_equal(a.g(), b.g())
    assert np.array_equal(a.moments(), b.moments())
    g.close()
