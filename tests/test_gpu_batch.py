"""GPU tests of several 2-D clouds stepped in one context (ab_batch_create).

Each cloud of a batch steps exactly as it does alone: the kernels are per
particle, the mass sums run per cloud in the same order, and each cloud keeps
its own MLS radius and table classes. So the checks are bit-identity:
  batch steps vs each member stepped alone      identical g and moments
  graph replays of a batch vs its eager steps   identical
  total mass                                    sum of the members' (1e-13)
"""
from __future__ import annotations

import numpy as np
import pytest

import alebgk4paper_amd as A
from helpers import setup

pytestmark = pytest.mark.gpu

SIZES = (10, 13, 12)


def _say(msg):
    print(msg, flush=True)


def _clouds(nv=8, order=2):
    return [setup("cavity2d", nx, nv, order=order)[1] for nx in SIZES]


def _dt():
    return 0.5 * (1e-6 / (max(SIZES) - 1)) / 1303.0


@pytest.mark.parametrize("prec", [A.PREC_FP64, A.PREC_FP64_TABLE])
def test_batch_steps_match_members_alone(prec):
    _say(f"[batch] cavity2d {SIZES} nv8, precision {A.PREC_NAMES[prec]}, 6 steps")
    clouds = _clouds()
    members = [A.Solver(cl, precision=prec) for cl in clouds]
    batch = A.Batch(members)
    dt = _dt()
    batch.steps(dt, 6)
    g, mom = batch.g(), batch.moments()
    for (a, b), m in zip(batch.ranges, members):
        m.steps(dt, 6)
        assert np.array_equal(g[a:b], m.g())
        assert np.array_equal(mom[a:b], m.moments())
    batch.close()


def test_batch_scatter_and_mass():
    _say(f"[batch] scatter and mass, cavity2d {SIZES} nv8")
    clouds = _clouds()
    members = [A.Solver(cl, precision=A.PREC_FP64_TABLE) for cl in clouds]
    batch = A.Batch(members)
    dt = _dt()
    batch.steps(dt, 3)
    batch.scatter()
    g = batch.g()
    total = 0.0
    for (a, b), m in zip(batch.ranges, members):
        assert np.array_equal(m.g(), g[a:b])
        total += m.total_mass()
    assert abs(batch.total_mass() - total) <= 1e-13 * total
    # the members carry on alone from the batch's state
    alone = A.Solver(clouds[1], precision=A.PREC_FP64_TABLE)
    alone.steps(dt, 5)
    members[1].steps(dt, 2)
    assert np.array_equal(members[1].g(), alone.g())
    batch.close()


def test_batch_graph_replays_match_eager():
    _say(f"[batch] graph replays, cavity2d {SIZES} nv8")
    dt = _dt()
    eager = A.Batch([A.Solver(cl, precision=A.PREC_FP64_TABLE) for cl in _clouds()])
    capt = A.Batch([A.Solver(cl, precision=A.PREC_FP64_TABLE) for cl in _clouds()])
    graph = A.Graph([capt], dt, 4)
    for _ in range(3):
        graph.launch()
        graph.synchronize()
        eager.steps(dt, 4)
        eager.synchronize()
        assert np.array_equal(capt.g(), eager.g())
    graph.close()


def test_batch_refusals():
    _say("[batch] refusals")
    c2 = setup("cavity2d", 10, 8, order=2)[1]
    c2b = setup("cavity2d", 10, 6, order=2)[1]
    c3 = setup("cavity3d", 6, 5, order=2)[1]
    s2, s2b, s3 = (A.Solver(c) for c in (c2, c2b, c3))
    with pytest.raises(A.ALEBGKError):
        A.Batch([s2, s3])                       # a 3-D cloud
    with pytest.raises(A.ALEBGKError):
        A.Batch([s2, s2b])                      # another velocity grid
    with pytest.raises(A.ALEBGKError):
        A.Batch([s2, A.Solver(c2, precision=A.PREC_FP64_TABLE)])   # another precision
    batch = A.Batch([s2])
    with pytest.raises(A.ALEBGKError, match="neighbour lists"):
        batch.rebuild_neighbours()
    with pytest.raises(A.ALEBGKError):
        batch.upload(c2)
    batch.close()


def test_batch_of_body_clouds_matches_alone():
    """Clouds around bodies (the reference harness's circle and square
    clouds, managed around their moved bodies on the device) batched with
    the direct FP64 transport: each steps as it does alone."""
    from conftest import load_fixture
    from test_gpu_cloud import device_body, solver_for
    _say("[batch] circle and square body clouds, FP64")
    fixtures = [load_fixture(n) for n in ("cloud_circle_n24_v6", "cloud_square_n24_v6")]

    def managed(fx):
        m = fx.meta
        sol = solver_for(fx)
        dom = (m["xleft"], m["xright"], m["ybottom"], m["ytop"], m["zfront"], m["zback"])
        sol.cloud_manage([device_body(fx["body"])], fx["bodyid_before"], m["dx"], m["minDist"],
                         dom)
        return sol

    members = [managed(fx) for fx in fixtures]
    alone = [managed(fx) for fx in fixtures]
    batch = A.Batch(members)
    batch.steps(1e-12, 4)
    g, mom = batch.g(), batch.moments()
    for (a, b), s in zip(batch.ranges, alone):
        s.steps(1e-12, 4)
        assert np.array_equal(g[a:b], s.g())
        assert np.array_equal(mom[a:b], s.moments())
    batch.close()
