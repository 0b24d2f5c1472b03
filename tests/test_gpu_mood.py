"""GPU: the reference's CPU step with its MOOD limiter (AB_LIMITER_MOOD,
runSolver, Solver/Solver.cpp:641-693) against that step itself, three steps
the reference's own functions ran from two opposing streams (fixtures of the
golden harness), and against the oracle's restatement.

Bars: the particles the transport check flags identical at every step; g and
the moments within 1e-9 of the reference after three steps (the reference
solves the MLS normal equations by Gauss elimination, this library by
Cholesky) and within 1e-10 of the oracle.
"""
from __future__ import annotations

import numpy as np
import pytest

import alebgk4paper_amd as A
import oracle_ctypes as O
from conftest import load_fixture
from helpers import rel_err, setup

pytestmark = pytest.mark.gpu

CASES = [("mood_cavity3d_n7_v6", "cavity3d", 7, 6), ("mood_cavity2d_n12_v8", "cavity2d", 12, 8)]


def _say(msg):
    print(msg, flush=True)


@pytest.mark.parametrize("fixture,name,nx,nv", CASES)
def test_mood_step_matches_reference_cpu_step(fixture, name, nx, nv):
    _say(f"[mood] {fixture}: 3 CPU-semantics steps vs the reference and the oracle")
    fx = load_fixture(fixture)
    m = fx.meta
    case, cl, ocl, neigh, nn = setup(name, nx, nv, order=2, perturb=False)
    cl.g = np.ascontiguousarray(fx["g0"])
    cl.mom = np.ascontiguousarray(fx["mom0"])
    sol = A.Solver(cl, precision=A.PREC_FP64, limiter=A.LIMITER_MOOD)
    st = O.State(ocl, fx["g0"], fx["mom0"])
    flags = fx["flags"]
    for k in range(flags.shape[0]):
        stats = sol.step(m["dt_transport"], stats=True)
        O.step_mood(ocl, st, m["dt_transport"], m["tau"], m["R"], m["radius"], 2)
        _, lim = sol.moments(with_limflag=True)
        mine = ((lim.astype(int) & 2) != 0).astype(np.uint8)
        _say(f"  step {k + 1}: {int(mine.sum())} flagged (reference {int(flags[k].sum())}), "
             f"positivity {stats['n_positivity']}, post-collision fails {stats['n_pad_fail']}")
        assert np.array_equal(mine, flags[k]), (k, np.nonzero(mine != flags[k]))
        assert stats["n_nad"] == int(flags[k].sum())
        assert np.array_equal(lim, st.limflag), k
    g = sol.g()
    mom = sol.moments()
    e_ref, e_orc = rel_err(g, fx["g1"]), rel_err(g, st.g)
    _say(f"  g vs reference {e_ref:.2e}, vs oracle {e_orc:.2e}")
    assert e_ref < 1e-9 and e_orc < 1e-10
    mom1 = fx["mom1"]
    for c in (0, 4):
        assert rel_err(mom[:, c], mom1[:, c]) < 1e-9, c
    assert np.abs(mom[:, 1:4] - mom1[:, 1:4]).max() < 1e-9 * np.abs(mom1[:, 1:4]).max()


def test_mood_refuses_table_precision_and_discrete_collision():
    case, cl, ocl, neigh, nn = setup("cavity3d", 6, 5, order=2, precision=A.PREC_MIXED)
    with pytest.raises(A.ALEBGKError):
        A.Solver(cl, precision=A.PREC_MIXED, limiter=A.LIMITER_MOOD)
    with pytest.raises(A.ALEBGKError):
        A.Solver(cl, precision=A.PREC_FP64, limiter=A.LIMITER_MOOD,
                 collision=A.COLLISION_DISCRETE)
