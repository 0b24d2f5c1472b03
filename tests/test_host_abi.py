"""CPU tests of the product's host side and of the C ABI's presence.

No GPU calls: the HIP library is only loaded and its exports checked against
include/alebgk_dev.h; the host library builds the reference's cases and is
compared bit for bit with the reference goldens."""
from __future__ import annotations

import ctypes as C
import os
import re

import numpy as np
import pytest

import alebgk4paper_amd as A
from conftest import ROOT, load_fixture


def _declared(header, prefix):
    txt = open(os.path.join(ROOT, "include", header)).read()
    return sorted(set(re.findall(r"\b(" + prefix + r"[a-z_0-9]+)\s*\(", txt)))


def test_hip_library_exports_every_declared_symbol():
    names = _declared("alebgk_dev.h", "ab_")
    assert len(names) >= 15
    lib = C.CDLL(A.HIP_LIB)
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert set(names) == set(A.HIP_EXPORTS)


def test_host_library_exports_every_declared_symbol():
    names = _declared("alebgk_host.h", "alh_")
    lib = C.CDLL(A.HOST_LIB)
    assert all(hasattr(lib, n) for n in names)
    assert set(names) == set(A.HOST_EXPORTS)


def test_case_table():
    """configureCase values (TestCases.cpp:113-228) and the overrides."""
    c = A.configure("cavity3d")
    assert (c.id, c.dim, c.nx, c.nv, c.dt, c.tau, c.rho0) == (2, 3, 31, 20, 1e-11, 3.7142e-10, 1.0)
    assert c.mls_order == 1                      # Types.hpp:148 default
    c = A.configure("6", nx=50)
    assert (c.name, c.nx, c.nz, c.nv, c.rho0, c.dt, c.lid_speed) == \
        ("cavity3d-cube", 50, 50, 10, 4.4, 1.5e-11, 10.0)
    c = A.configure("cavity2d-circle")
    assert (c.dim, c.nx, c.nv, c.vmax, c.tau) == (2, 60, 30, 1500, 3.37e-11)
    with pytest.raises(A.ALEBGKError):
        A.configure("cavity4d")
    with pytest.raises(A.ALEBGKError):
        A.configure("7")


@pytest.mark.parametrize("name,case", [("cavity3d_n6_v5", "cavity3d"),
                                       ("cavity3d_n9_v3", "cavity3d"),
                                       ("cavity2d_n10_v8", "cavity2d")])
def test_host_lattice_matches_reference(name, case):
    fx = load_fixture(name)
    m = fx.meta
    cs = A.configure(case, nx=m["Nx"], nv=m["Nv"])
    cl = A.build_lattice(cs)
    for k in ("x", "y", "z"):
        assert np.array_equal(getattr(cl, k), fx[k]), k
    assert np.array_equal(cl.boundary, fx["boundary"])
    b = cl.boundary == 1
    assert np.array_equal(cl.wall[b], fx["wall"][b])
    assert np.array_equal(cl.g, fx["g_ic"])
    cfg = cl.cfg
    assert (cfg.hbox, cfg.vbox, cfg.radius) == (m["hBox"], m["vBox"], m["radius"])
    assert (cfg.nbhbox, cfg.nbvbox) == (m["nbhBox"], m["nbvBox"])
    if m["dim"] == 3:
        assert (cfg.wbox, cfg.nbwbox) == (m["wBox"], m["nbwBox"])


def test_host_lattice_rejects_body_cases():
    with pytest.raises(A.ALEBGKError):
        A.build_lattice(A.configure("cavity3d-cube", nx=8, nv=4))


def test_solver_fails_loudly_without_gpu():
    """No CPU fallback: without a device the context cannot be created."""
    try:
        import torch  # noqa: F401
        if torch.cuda.is_available():
            pytest.skip("a GPU is present")
    except ImportError:
        pass
    cl = A.build_lattice(A.configure("cavity3d", nx=6, nv=4))
    with pytest.raises(A.ALEBGKError):
        A.Solver(cl)


def test_check_config_refuses_table_span_beyond_2gib():
    """The 3-D table transport addresses one z block (16 lattice planes) of
    stored g with 32-bit byte offsets. cavity3d 60^3, Nv=20: 16 planes of
    60*60*8000 values are 1.84 GB in FP32 (accepted) and 3.69 GB in FP64
    (refused before any device work); the FP64 reference formulation has no
    such block and is accepted."""
    cl = A.build_lattice(A.configure("cavity3d", nx=6, nv=4), mls_order=2)
    cfg = cl.cfg
    cfg.nx = cfg.ny = cfg.nz = 60
    cfg.n_particles = 60 ** 3
    cfg.nv = 20
    for prec, want in ((A.PREC_MIXED, 0), (A.PREC_FP64_TABLE, 4), (A.PREC_FP64, 0)):
        cfg.precision = prec
        rc, msg = A.check_config(cfg)
        assert rc == want, (prec, rc, msg)
        assert (rc == 0) == (msg == ""), msg
    cfg.precision = A.PREC_FP64_TABLE
    assert "2 GiB" in A.check_config(cfg)[1]
    # the largest FP32 lattice at Nv=20 that fits: 16 * n^2 * 8000 * 4 < 2^31
    cfg.precision = A.PREC_MIXED
    cfg.nx = cfg.ny = cfg.nz = 64
    cfg.n_particles = 64 ** 3
    assert A.check_config(cfg)[0] == 0
    cfg.nx = cfg.ny = cfg.nz = 65
    cfg.n_particles = 65 ** 3
    assert A.check_config(cfg)[0] == 4
    # argument ranges
    cfg.nx = cfg.ny = cfg.nz = 6
    cfg.n_particles = 216
    cfg.nv = 49
    assert A.check_config(cfg)[0] == 1
    cfg.nv = 8
    cfg.dim = 2
    cfg.collision = A.COLLISION_DISCRETE
    assert A.check_config(cfg)[0] == 4


@pytest.mark.parametrize("name,case,nx,nv", [("body_cube_n12_v4", "cavity3d-cube", 12, 4),
                                             ("body_sphere_n12_v4", "cavity3d-sphere", 12, 4),
                                             ("body_circle_n24_v6", "cavity2d-circle", 24, 6),
                                             ("body_square_n24_v6", "cavity2d-square", 24, 6)])
def test_host_body_case_matches_reference(name, case, nx, nv):
    """makeBodies + embedRigidBodies + initWallData + applyLidProfile +
    applyIC for the four body cases at body scale 2: the same particles in
    the same order, the same flags, body ids, wall state (the parabolic lid
    and the body normals included), initial state and body, bit for bit."""
    fx = load_fixture(name)
    cs = A.configure(case, nx=nx, nv=nv)
    cl = A.build_case(cs, body_scale=float(fx["params"][0]))
    for k in ("x", "y", "z"):
        assert np.array_equal(getattr(cl, k), fx[k]), k
    assert np.array_equal(cl.boundary, fx["boundary"])
    assert np.array_equal(cl.body_id, fx["bodyid"])
    assert np.array_equal(cl.wall, fx["wall"])
    assert np.array_equal(cl.mom, fx["mom_ic"])
    rows = fx["g_ic_rows"]
    assert np.array_equal(cl.g[0], rows[0]) and np.array_equal(cl.g[-1], rows[1])
    assert np.all(cl.g == cl.g[0])
    b = cl.bodies[0]
    r = fx["body"]
    assert (b.shape, b.npts) == (int(r[0]), int(r[1]))
    assert (b.mass, b.T, b.size) == (r[2], r[3], r[4])
    assert list(b.c) == list(r[5:8]) and list(b.o) == list(r[8:10])
    assert list(b.rm) == list(r[10:19])
    m = fx.meta
    assert cl.geo[0] == m["dx"] and cl.geo[3] == m["minDist"]
    assert cl.cfg.radius == m["radius"] and cl.cfg.n_particles == m["N"]
    # the lid: a parabolic profile peaking at 10 m/s on y = ytop
    lid = (cl.boundary == 1) & (cl.body_id < 0) & (np.abs(cl.y - 1e-6) <= 1e-13)
    assert lid.sum() > 0 and cl.wall[lid, 3].max() <= 10.0


def test_lid_parabolic_oracle_matches_reference():
    """orc_lid_parabolic against applyLidProfile's output in the body-case
    fixtures: the oracle, given the lattice and the uniform wall state,
    writes exactly the reference's lid velocities, and touches neither the
    body surface nor any other wall."""
    import oracle_ctypes as O
    for name, nx in (("body_cube_n12_v4", 12), ("body_circle_n24_v6", 24)):
        fx = load_fixture(name)
        m = fx.meta
        dim = m["dim"]
        geo, *_ = O.lattice(dim, nx, (m["xleft"], m["xright"], m["ytop"], m["ybottom"],
                                      m["zfront"], m["zback"]), Ny=nx, Nz=nx if dim == 3 else 1)
        want = fx["wall"][:, 3]
        wux = np.where((fx["boundary"] == 1) & (fx["bodyid"] < 0) &
                       (np.abs(fx["y"] - m["ytop"]) <= 1e-13), 0.0, want)
        got = np.ascontiguousarray(wux.copy())
        O.lid_parabolic(geo, fx["x"], fx["y"], fx["boundary"], fx["bodyid"], 10.0, got)
        assert np.array_equal(got, want), name
        assert np.count_nonzero(got != wux) > 0
