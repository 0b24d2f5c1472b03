"""GPU: cloud management around a moved body on the device (ab_cloud_manage:
manageRemove, mergeCloseParticles, manageRefill with the neighbour
rebuilds, RigidBody/CloudManagement.cpp:17-275, Solver/Solver.cpp:379-398)
against the reference's own functions (golden-harness fixtures after a body
move) and the oracle.

Bars: identical counts, positions, boundary flags and order (the source map
equals the oracle's); g and the moments of kept and merged particles
identical; a refilled particle's moments within 1e-13 (its weighted sums are
a block reduction, the reference's are sequential) and its Maxwellian row
within 1e-13; the wall state carried along; a step runs on the new cloud.
"""
from __future__ import annotations

import numpy as np
import pytest

import alebgk4paper_amd as A
import oracle_ctypes as O
from conftest import load_fixture
from helpers import rel_err

pytestmark = pytest.mark.gpu

CASES = ["cloud_sphere_n12_v3", "cloud_cube_n12_v3", "cloud_circle_n24_v6",
         "cloud_square_n24_v6"]


def _say(msg):
    print(msg, flush=True)


def solver_for(fx):
    m = fx.meta
    N = m["N"]
    cfg = A.ABConfig()
    cfg.dim, cfg.n_particles, cfg.nv = m["dim"], N, m["Nv"]
    cfg.vmin, cfg.vmax, cfg.R, cfg.tau = m["vmin"], m["vmax"], m["R"], m["tau"]
    cfg.radius, cfg.mls_order, cfg.precision = m["radius"], 2, A.PREC_FP64
    cfg.xleft, cfg.ybottom, cfg.zfront = m["xleft"], m["ybottom"], m["zfront"]
    cfg.hbox, cfg.vbox, cfg.wbox = m["hBox"], m["vBox"], m["wBox"]
    cfg.nbhbox, cfg.nbvbox, cfg.nbwbox = m["nbhBox"], m["nbvBox"], m["nbwBox"]
    cl = A.Cloud(cfg, fx["x_before"], fx["y_before"], fx["z_before"], fx["boundary_before"],
                 fx["wall_before"], fx["mom_before"], fx["g_before"])
    return A.Solver(cl)


def device_body(b):
    return A.body(int(b[0]), b[1:4], b[4], o=b[5:7], rm=b[7:16])


@pytest.mark.parametrize("name", CASES)
def test_cloud_management_matches_reference(name):
    fx = load_fixture(name)
    m = fx.meta
    _say(f"[cloud] {name}: N={m['N']}, expected counts {tuple(fx['counts'])}")
    sol = solver_for(fx)
    dom = (m["xleft"], m["xright"], m["ybottom"], m["ytop"], m["zfront"], m["zback"])
    src, counts = sol.cloud_manage([device_body(fx["body"])], fx["bodyid_before"], m["dx"],
                                   m["minDist"], dom)
    _say(f"  device counts {counts}, N {sol.N}")
    assert counts == tuple(int(v) for v in fx["counts"])
    ref = O.cloud_manage(m, *[fx[k + "_before"] for k in ("x", "y", "z", "boundary", "bodyid",
                                                           "g", "mom")],
                         [O.body_from_vector(fx["body"])])
    assert np.array_equal(src, ref["src"])
    p = sol.particles()
    for k in ("x", "y", "z"):
        assert np.array_equal(p[k], fx[k + "_after"]), k
    assert np.array_equal(p["boundary"], fx["boundary_after"])
    kept = src >= 0
    assert np.array_equal(p["wall"][kept], fx["wall_before"][src[kept]])
    assert np.all(p["wall"][~kept] == 0.0)
    g, mom = sol.g(), sol.moments()
    ga, ma = fx["g_after"], fx["mom_after"]
    assert np.array_equal(g[kept], ga[kept])
    assert np.array_equal(mom[kept], ma[kept])
    if (~kept).any():
        e_m = rel_err(mom[~kept], ma[~kept])
        e_g = rel_err(g[~kept], ga[~kept])
        _say(f"  refilled: moments {e_m:.1e}, g {e_g:.1e}")
        assert e_m < 1e-13 and e_g < 1e-13
    # the neighbour lists follow the new cloud, and a step runs on it
    nb, nn = sol.neighbours()
    rneigh, rnn, _ = O.neighbours(O.geom_from_meta(m), p["x"], p["y"], p["z"])
    assert np.array_equal(nn, rnn)
    for i in range(sol.N):
        assert np.array_equal(nb[i, : nn[i]], rneigh[i, : rnn[i]])
    sol.step(1e-12)
    assert np.isfinite(sol.moments()).all()
    sol.close()


def test_cloud_management_without_bodies_keeps_the_cloud():
    fx = load_fixture("cloud_circle_n24_v6")
    m = fx.meta
    sol = solver_for(fx)
    dom = (m["xleft"], m["xright"], m["ybottom"], m["ytop"], m["zfront"], m["zback"])
    src, counts = sol.cloud_manage([], fx["bodyid_before"], m["dx"], m["minDist"], dom)
    assert counts == (0, 0, 0) and np.array_equal(src, np.arange(m["N"]))
    sol.close()


def test_cloud_management_refuses_the_table_pipeline():
    case = A.configure("cavity3d", nx=6, nv=4, mls_order=2)
    cl = A.build_lattice(case, precision=A.PREC_MIXED, mls_order=2)
    sol = A.Solver(cl, precision=A.PREC_MIXED)
    with pytest.raises(A.ALEBGKError):
        sol.cloud_manage([A.body(A.BODY_SPHERE, (5e-7, 5e-7, 5e-7), 1e-7)],
                         np.full(cl.N, -1, np.int32), 1e-7, 0.0, case.dom)
    sol.close()
