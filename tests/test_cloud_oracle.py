"""The oracle's restatement of the cloud management of the moving-body step
(manageRemove, mergeCloseParticles, manageRefill:
RigidBody/CloudManagement.cpp:17-275, sequenced as Solver/Solver.cpp:379-398)
against the reference's own functions: fixtures of the golden harness after
a body move (a sphere and a turned cube in 3-D, a circle and a square in
2-D; the interior jittered so that pairs merge).

Bar: identical counts, positions, flags and body ids; g and the moments
identical (the merge averages in the reference's expression; the refill
sums in the reference's order).

The reference's 3-D refill adds nothing in these cavities: its site range
in z runs from zfront to zback, and the 3-D cavity has zback < zfront
(CloudManagement.cpp:162, :168-170), so the range is empty. The 3-D fixtures
exercise remove and merge, the 2-D ones all three.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

import oracle_ctypes as O
from conftest import load_fixture

CASES = ["cloud_sphere_n12_v3", "cloud_cube_n12_v3", "cloud_circle_n24_v6",
         "cloud_square_n24_v6"]


def before(fx):
    return [fx[k + "_before"] for k in ("x", "y", "z", "boundary", "bodyid", "g", "mom")]


@pytest.mark.parametrize("name", CASES)
def test_oracle_cloud_management_matches_reference(name):
    fx = load_fixture(name)
    m = fx.meta
    x, y, z, b, bid, g, mom = before(fx)
    out = O.cloud_manage(m, x, y, z, b, bid, g, mom, [O.body_from_vector(fx["body"])])
    assert tuple(out["counts"]) == tuple(int(v) for v in fx["counts"])
    for k in ("x", "y", "z"):
        assert np.array_equal(out[k], fx[k + "_after"]), k
    assert np.array_equal(out["boundary"], fx["boundary_after"])
    assert np.array_equal(out["bodyid"], fx["bodyid_after"])
    assert np.array_equal(out["g"], fx["g_after"])
    assert np.array_equal(out["mom"], fx["mom_after"])


def test_fixtures_exercise_every_stage():
    """The 3-D cases remove and merge; the 2-D ones also refill."""
    c = {n: tuple(int(v) for v in load_fixture(n)["counts"]) for n in CASES}
    assert all(r > 0 and mg > 0 for r, mg, _ in c.values())
    assert c["cloud_circle_n24_v6"][2] > 0 and c["cloud_square_n24_v6"][2] > 0


def test_inside_tests_match_the_reference_shapes():
    """orc_in_body against the shapes' closed forms on random points: a
    point is inside a sphere/circle within the radius, inside a square/cube
    within half the side along the body axes."""
    rng = np.random.default_rng(5)
    pts = rng.uniform(-1.0, 1.0, size=(2000, 3))
    L = O.lib()
    L.orc_in_body.restype = C.c_int
    cases = [
        (np.array([3, 0.1, -0.2, 0.05, 0.6] + [1, 0] + [1, 0, 0, 0, 1, 0, 0, 0, 1.0]),
         lambda p: np.sum((p - [0.1, -0.2, 0.05]) ** 2, 1) <= 0.36, 1),
        (np.array([2, 0.1, -0.2, 0.0, 0.6] + [1, 0] + [1, 0, 0, 0, 1, 0, 0, 0, 1.0]),
         lambda p: np.sum((p[:, :2] - [0.1, -0.2]) ** 2, 1) <= 0.36, 0),
    ]
    a = 0.4
    c, s = np.cos(a), np.sin(a)
    Rm = np.array([[c, -s, 0], [s, c, 0], [0, 0, 1.0]])
    cases.append((np.concatenate([[4, 0.1, -0.2, 0.05, 0.8, 1, 0], Rm.ravel()]),
                  lambda p: np.all(np.abs((p - [0.1, -0.2, 0.05]) @ Rm) <= 0.4, 1), 1))
    o = np.array([c, s])
    cases.append((np.concatenate([[1, 0.1, -0.2, 0.0, 0.8], o, np.eye(3).ravel()]),
                  lambda p: (np.abs((p[:, 0] - 0.1) * o[0] + (p[:, 1] + 0.2) * o[1]) <= 0.4)
                  & (np.abs((p[:, 0] - 0.1) * o[1] - (p[:, 1] + 0.2) * o[0]) <= 0.4), 0))
    for vec, inside, d3 in cases:
        bo = O.body_from_vector(vec)
        got = np.array([L.orc_in_body(C.byref(bo), C.c_double(p[0]), C.c_double(p[1]),
                                      C.c_double(p[2]), d3) for p in pts], bool)
        assert np.array_equal(got, inside(pts)), int(vec[0])
