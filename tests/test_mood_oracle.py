"""The oracle's restatement of the reference's CPU step with its MOOD limiter
(orc_step_mood, Solver/Solver.cpp:641-693) against that step itself: the
fixtures are three steps the reference's own functions ran (the golden
harness, tests/golden/make_golden.sh) from two opposing streams, a start that
makes the transport check flag particles and re-transport them.

Bars: the flagged set of every step identical; g and the moments after three
steps within 1e-9 (the reference solves the MLS normal equations by Gauss
elimination, the oracle by Cholesky; they agree to round-off of these
stencils, and a flag decided the other way would show as an O(1) error of a
particle's row).
"""
from __future__ import annotations

import numpy as np
import pytest

import oracle_ctypes as O
from conftest import load_fixture
from helpers import rel_err, setup

CASES = [("mood_cavity3d_n7_v6", "cavity3d", 7, 6), ("mood_cavity2d_n12_v8", "cavity2d", 12, 8)]


@pytest.mark.parametrize("fixture,name,nx,nv", CASES)
def test_oracle_mood_step_matches_reference_cpu_step(fixture, name, nx, nv):
    fx = load_fixture(fixture)
    m = fx.meta
    case, cl, ocl, neigh, nn = setup(name, nx, nv, order=2, perturb=False)
    assert cl.N == m["N"]
    g0, mom0 = fx["g0"], fx["mom0"]
    st = O.State(ocl, g0, mom0)
    flags = fx["flags"]
    nsteps = flags.shape[0]
    assert flags[0].sum() > 0                  # the start makes the check fire
    for k in range(nsteps):
        O.step_mood(ocl, st, m["dt_transport"], m["tau"], m["R"], m["radius"], 2)
        mine = ((st.limflag.astype(int) & 2) != 0).astype(np.uint8)
        assert np.array_equal(mine, flags[k]), (k, np.nonzero(mine != flags[k]))
    assert rel_err(st.g, fx["g1"]) < 1e-9
    mom1 = fx["mom1"]
    for c in (0, 4):
        assert rel_err(st.mom[:, c], mom1[:, c]) < 1e-9, c
    u = st.mom[:, 1:4]
    assert np.abs(u - mom1[:, 1:4]).max() < 1e-9 * np.abs(mom1[:, 1:4]).max()


def test_oracle_mood_differs_from_the_gpu_semantics():
    """The two limiter semantics differ on the same start (positivity to
    zero vs fallback, NAD re-transport vs rescale, one mass rescale vs two),
    so the MOOD fixture is not passed by the GPU-path step by accident."""
    fx = load_fixture("mood_cavity3d_n7_v6")
    m = fx.meta
    case, cl, ocl, neigh, nn = setup("cavity3d", 7, 6, order=2, perturb=False)
    st = O.State(ocl, fx["g0"], fx["mom0"])
    for _ in range(fx["flags"].shape[0]):
        O.step(ocl, st, m["dt_transport"], m["tau"], m["R"], m["radius"], 2)
    assert rel_err(st.g, fx["g1"]) > 1e-6
