"""GPU parity: the HIP library (through its C ABI) against the CPU oracle.

Tolerances (stated per check):
  neighbour lists                 bit-exact
  FP64 transport                  1e-12 of each particle's largest |g|
  FP64 step (g, rho, u, T)        1e-10 relative (north star: FP64 moments
                                  to 1e-10 of the oracle)
  FP64_TABLE vs FP64              1e-10 (same arithmetic, coefficients
                                  factorised once)
  MIXED vs FP64                   1e-5 relative on rho, u, T (north star)
  body loads                      1e-12 relative to the reference golden
"""
from __future__ import annotations

import numpy as np
import pytest

import alebgk4paper_amd as A
import oracle_ctypes as O
from conftest import load_fixture
from helpers import moment_errors, rel_err, setup

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


# ------------------------------------------------------------- neighbours ----

@pytest.mark.parametrize("name,case", [("cavity3d_n6_v5", "cavity3d"),
                                       ("cavity3d_n9_v3", "cavity3d"),
                                       ("cavity2d_n10_v8", "cavity2d")])
def test_device_neighbours_match_reference(name, case):
    fx = load_fixture(name)
    m = fx.meta
    _say(f"[neigh] {name}")
    cl = A.build_lattice(A.configure(case, nx=m["Nx"], nv=m["Nv"]))
    sol = A.Solver(cl)                       # no lists given: device rebuild
    nb, nn = sol.neighbours()
    assert np.array_equal(nn, fx["nneigh"])
    assert np.array_equal(nb, fx.neigh_padded())


def test_device_neighbours_match_oracle_metric_size():
    _say("[neigh] cavity3d 40^3")
    case, cl, ocl, neigh, nn = setup("cavity3d", 40, 3, perturb=False)
    sol = A.Solver(cl)
    nb, dnn = sol.neighbours()
    assert np.array_equal(dnn, nn)
    assert np.array_equal(nb, neigh)
    assert nn.max() == 80


@pytest.mark.parametrize("case,nx", [("cavity3d", 16), ("cavity2d", 40)])
def test_device_neighbours_jittered_permuted_cloud(case, nx):
    """A cloud that is no lattice: positions jittered by up to 0.45 dx and
    the particles in a random order, so every voxel's scattered segment must
    be sorted and the candidates cross voxel rows unevenly."""
    from helpers import geom_of
    _say(f"[neigh] {case} {nx}, jittered and permuted")
    cs = A.configure(case, nx=nx, nv=4)
    cl = A.build_lattice(cs)
    rng = np.random.default_rng(11)
    dx = cl.cfg.radius / 2.5                 # the MLS radius is 2.5 dx
    perm = rng.permutation(cl.N)
    for k in ("x", "y") + (("z",) if cl.cfg.dim == 3 else ()):
        a = getattr(cl, k) + 0.45 * dx * rng.uniform(-1, 1, cl.N)
        setattr(cl, k, np.ascontiguousarray(a[perm]))
    neigh, nn, _ = O.neighbours(geom_of(cl.cfg, cs), cl.x, cl.y, cl.z)
    sol = A.Solver(cl)
    nb, dnn = sol.neighbours()
    assert np.array_equal(dnn, nn)
    assert np.array_equal(nb, neigh)
    sol.rebuild_neighbours()                 # again on the same cloud
    nb2, dnn2 = sol.neighbours()
    assert np.array_equal(nb2, nb) and np.array_equal(dnn2, dnn)


def test_device_neighbours_dense_cloud_caps_and_unstaged_voxels():
    """A 16^3 cloud squeezed into a third of the box along each axis (same
    voxel grid and radius): about 27 times the lattice density, so every
    particle has more than 128 candidates in range (the list is capped at
    the first 128 in scan order) and voxels see more candidates than the
    search stages in LDS (those read them from memory in the same order)."""
    from helpers import geom_of
    _say("[neigh] dense cloud: caps and unstaged voxels")
    cs = A.configure("cavity3d", nx=16, nv=4)
    cl = A.build_lattice(cs)
    rng = np.random.default_rng(3)
    dx = cl.cfg.radius / 2.5
    for k in ("x", "y", "z"):
        a = getattr(cl, k)
        lo = a.min()
        setattr(cl, k, np.ascontiguousarray(lo + (a - lo) / 3.0 + 0.01 * dx * rng.uniform(-1, 1, cl.N)))
    neigh, nn, _ = O.neighbours(geom_of(cl.cfg, cs), cl.x, cl.y, cl.z)
    assert nn.max() == 128 and (nn == 128).mean() > 0.5
    sol = A.Solver(cl)
    nb, dnn = sol.neighbours()
    assert np.array_equal(dnn, nn)
    assert np.array_equal(nb, neigh)


# -------------------------------------------------------------- transport ----

@pytest.mark.parametrize("nx,nv,order", [(6, 5, 1), (6, 5, 2), (9, 4, 2), (12, 6, 2)])
def test_fp64_transport_matches_oracle(nx, nv, order):
    _say(f"[transport] cavity3d nx={nx} nv={nv} order={order}")
    case, cl, ocl, neigh, nn = setup("cavity3d", nx, nv, order=order)
    cl.neigh, cl.nneigh = neigh, nn
    dx = 1e-6 / (nx - 1)
    dt = 0.5 * dx / 1303.0
    sol = A.Solver(cl, precision=A.PREC_FP64)
    sol.transport_only(dt)
    gt = sol.gt()
    want, fl, _ = O.transport(ocl, cl.g, dt, cl.cfg.radius, order, flags=True)
    scale = np.abs(cl.g).max(axis=1, keepdims=True)
    err = np.abs(gt - want) / scale
    assert err.max() < 1e-12, err.max()


@pytest.mark.parametrize("nx,nv,order", [(10, 6, 1), (10, 6, 2), (12, 8, 1), (12, 8, 2)])
def test_fp64_transport_2d_matches_oracle(nx, nv, order):
    _say(f"[transport2d] cavity2d nx={nx} nv={nv} order={order}")
    case, cl, ocl, neigh, nn = setup("cavity2d", nx, nv, order=order)
    dx = 1e-6 / (nx - 1)
    dt = 0.5 * dx / 1303.0
    sol = A.Solver(cl, precision=A.PREC_FP64)
    sol.transport_only(dt)
    gt = sol.gt()
    want = O.transport(ocl, cl.g, dt, cl.cfg.radius, order)
    scale = np.abs(cl.g).max(axis=1, keepdims=True)
    assert (np.abs(gt - want) / scale).max() < 1e-12


# ------------------------------------------------------------------ steps ----

@pytest.mark.parametrize("name,nx,nv,order", [("cavity3d", 8, 5, 2), ("cavity3d", 10, 6, 1),
                                              ("cavity2d", 12, 8, 2)])
def test_fp64_step_matches_oracle(name, nx, nv, order):
    _say(f"[step] {name} nx={nx} nv={nv} order={order}")
    case, cl, ocl, neigh, nn = setup(name, nx, nv, order=order)
    dx = 1e-6 / (nx - 1)
    dt = 0.5 * dx / 1303.0
    sol = A.Solver(cl, precision=A.PREC_FP64)
    st = O.State(ocl, cl.g, cl.mom)
    for k in range(3):
        stats = sol.step(dt, stats=True)
        m_old = O.step(ocl, st, dt, case.tau, case.R, cl.cfg.radius, order)
        assert abs(stats["mass_old"] - m_old) <= 1e-13 * m_old
    g = sol.g()
    mom, lim = sol.moments(with_limflag=True)
    assert rel_err(g, st.g) < 1e-10
    om = st.mom
    cols = [0, 1, 2, 3, 4] if cl.cfg.dim == 3 else [0, 1, 2, 4]
    for c in cols:
        assert rel_err(mom[:, c], om[:, c], floor=1e-300) < 1e-10, c
    assert np.array_equal(lim, st.limflag)
    # the step's report counts the same particles the oracle flags
    fl = st.limflag.astype(np.int64)
    assert stats["n_positivity"] == int(np.count_nonzero(fl & 1))
    assert stats["n_nad"] == int(np.count_nonzero(fl & 2))


def test_fp64_step_conserves_mass_and_is_deterministic():
    _say("[step] determinism cavity3d 10^3 nv6")
    case, cl, ocl, neigh, nn = setup("cavity3d", 10, 6, order=2)
    dt = 0.5 * (1e-6 / 9) / 1303.0
    runs = []
    for _ in range(2):
        sol = A.Solver(cl, precision=A.PREC_FP64)
        m0 = sol.total_mass()
        sol.steps(dt, 4)
        runs.append(sol.g())
        assert abs(sol.total_mass() - m0) < 1e-13 * m0
        sol.close()
    assert np.array_equal(runs[0], runs[1])


# ------------------------------------------------------- table pipelines ----

def test_fp64_table_matches_fp64_direct():
    _say("[table] fp64-table vs fp64, cavity3d 10^3 nv6")
    case, cl, ocl, neigh, nn = setup("cavity3d", 10, 6, order=2)
    dt = 0.5 * (1e-6 / 9) / 1303.0
    ref = A.Solver(cl, precision=A.PREC_FP64)
    tab = A.Solver(cl, precision=A.PREC_FP64_TABLE)
    for _ in range(3):
        ref.step(dt)
        tab.step(dt)
    assert rel_err(tab.g(), ref.g()) < 1e-10
    mr, mt = ref.moments(), tab.moments()
    for c in range(5):
        assert rel_err(mt[:, c], mr[:, c], floor=1e-300) < 1e-10, c
    assert abs(tab.total_mass() - ref.total_mass()) < 1e-12 * ref.total_mass()


@pytest.mark.parametrize("nx,nv,nsteps", [(10, 6, 20), (16, 8, 10)])
def test_mixed_moments_within_1e5_of_fp64(nx, nv, nsteps):
    _say(f"[mixed] vs fp64 cavity3d nx={nx} nv={nv} steps={nsteps}")
    case, cl, ocl, neigh, nn = setup("cavity3d", nx, nv, order=2)
    dt = case.dt
    ref = A.Solver(cl, precision=A.PREC_FP64)
    mix = A.Solver(cl, precision=A.PREC_MIXED)
    for _ in range(nsteps):
        ref.step(dt)
        mix.step(dt)
    inner = cl.boundary == 0
    e = moment_errors(mix.moments(), ref.moments(), inner)
    _say(f"  errors {e}")
    assert max(e.values()) < 1e-5, e
    assert abs(mix.total_mass() - ref.total_mass()) < 1e-9 * ref.total_mass()


def test_mixed_metric_size_moment_error():
    """The benchmark configuration (cavity3d 40^3, Nv 10, quadratic, lid
    1 m/s): after 50 steps the mixed path's rho, u and T agree with the FP64
    reference formulation to 1e-5 relative (|du| relative to max |u|)."""
    _say("[mixed] vs fp64 cavity3d 40^3 nv10, 50 steps")
    case = A.configure("cavity3d", nx=40, nv=10, mls_order=2)
    cl = A.build_lattice(case, mls_order=2)
    ref = A.Solver(cl, precision=A.PREC_FP64)
    mix = A.Solver(cl, precision=A.PREC_MIXED)
    ref.steps(case.dt, 50)
    mix.steps(case.dt, 50)
    e = moment_errors(mix.moments(), ref.moments(), cl.boundary == 0)
    _say(f"  errors {e}")
    assert max(e.values()) < 1e-5, e


def test_mixed_metric_size_properties():
    """cavity3d 40^3, Nv 10, quadratic: the benchmark configuration. Mass is
    restored to the FP64 sum, the field stays finite and positive, and the
    uniform state at rest with resting walls stays uniform."""
    _say("[mixed] cavity3d 40^3 nv10 properties")
    case = A.configure("cavity3d", nx=40, nv=10, mls_order=2)
    cl = A.build_lattice(case, precision=A.PREC_MIXED, mls_order=2)
    sol = A.Solver(cl)
    ref = A.Solver(cl, precision=A.PREC_FP64_TABLE)
    m0 = sol.total_mass()
    npos = []
    for _ in range(5):
        stats = sol.step(case.dt, stats=True)
        rst = ref.step(case.dt, stats=True)
        npos.append((stats["n_positivity"], rst["n_positivity"], stats["n_nad"], rst["n_nad"]))
    _say(f"  positivity / NAD counts (mixed, fp64-table) per step: {npos}")
    assert abs(sol.total_mass() - m0) < 1e-12 * m0
    g = sol.g()
    assert np.all(np.isfinite(g))
    mom = sol.moments()
    assert np.all(mom[:, 0] > 0) and np.all(np.isfinite(mom))
    # from rest the NAD clamp fires on most particles in the first step and
    # on the lid's neighbourhood after (54,872 then 15,324, 13,080, 684, 496
    # measured), the positivity repair on none; FP32 storage may flip a
    # borderline particle, no more
    for pm, pr, nm, nr in npos:
        assert nr > 0 and abs(pm - pr) <= 2 and abs(nm - nr) <= 2


# ------------------------------------------------------------- body loads ----

def test_body_loads_match_reference(cube):
    _say("[body] cube loads")
    m = cube.meta
    surf = cube["surf_idx"]
    N = m["N"]
    g = np.zeros((N, m["gStride"]))
    g[surf] = cube["g_surf"]
    wall = cube["wall"].copy()
    wall[surf, 3:6] = cube["surf_wU"]
    cfg = A.ABConfig()
    cfg.dim, cfg.n_particles, cfg.nv = 3, N, m["Nv"]
    cfg.vmin, cfg.vmax, cfg.R, cfg.tau = m["vmin"], m["vmax"], m["R"], m["tau"]
    cfg.radius, cfg.mls_order, cfg.precision = m["radius"], 1, A.PREC_FP64
    mom = np.tile([1.0, 0, 0, 0, 270.0], (N, 1))
    cl = A.Cloud(cfg, cube["x"], cube["y"], cube["z"], cube["boundary"], wall, mom, g,
                 neigh=cube.neigh_padded(), nneigh=cube["nneigh"])
    sol = A.Solver(cl)
    b = cube["body"]
    out = sol.body_loads(surf, np.zeros(len(surf), np.int32), b[None, :3], [b[3]])
    ref = cube["loads"]
    np.testing.assert_allclose(out[0], ref, rtol=1e-11, atol=1e-12 * np.abs(ref).max())
    # the same loads for a trial wall velocity given explicitly (Heun corrector)
    out2 = sol.body_loads(surf, np.zeros(len(surf), np.int32), b[None, :3], [b[3]],
                          wu=cube["surf_wU"])
    np.testing.assert_allclose(out2[0], ref, rtol=1e-11, atol=1e-12 * np.abs(ref).max())


# ------------------------------------------------------------- wall state ----

@pytest.mark.parametrize("prec", [A.PREC_FP64, A.PREC_FP64_TABLE])
def test_set_wall_state_matches_upload(prec):
    """ab_set_wall_state (cudaSyncBodies' uploads) changes positions' wall
    state on the device: the lid particles' wall velocity tripled and their
    mesh velocity set through it step exactly like a cloud uploaded with
    those values, and ab_download_particles returns them."""
    _say(f"[wall] set_wall_state, precision {prec}")
    case, cl, ocl, neigh, nn = setup("cavity3d", 9, 5, order=2)
    dt = 0.5 * (1e-6 / 8) / 1303.0
    lid = np.nonzero(np.abs(cl.z - case.dom[4]) < 1e-13)[0]
    wu = cl.wall[lid, 3:6] * 3.0
    wn = cl.wall[lid, 0:3].copy()
    a = A.Solver(cl, precision=prec)
    a.set_wall_state(lid, wn=wn, wu=wu)
    cl2 = A.Cloud(cl.cfg, cl.x, cl.y, cl.z, cl.boundary, cl.wall.copy(), cl.mom, cl.g)
    cl2.wall[lid, 3:6] = wu
    b = A.Solver(cl2, precision=prec)
    p = a.particles()
    assert np.array_equal(p["wall"][lid, 3:6], wu)
    assert np.array_equal(p["wall"], b.particles()["wall"])
    for _ in range(2):
        a.step(dt)
        b.step(dt)
    assert np.array_equal(a.g(), b.g())
    # positions and mesh velocity of a subset, read back exactly
    sub = lid[:5]
    xyz = np.stack([cl.x[sub], cl.y[sub], cl.z[sub]], axis=1)
    f = A.Solver(cl, precision=A.PREC_FP64)
    f.set_wall_state(sub, xyz=xyz + 1e-9, mv=np.ones((5, 3)))
    q = f.particles()
    assert np.array_equal(q["x"][sub], xyz[:, 0] + 1e-9)
    assert np.array_equal(np.delete(q["x"], sub), np.delete(cl.x, sub))


def test_body_loads_on_table_storage():
    """ab_body_loads reads the table pipelines' storage g = L (M0 + d): after
    one step (L != 1) the loads on the lid particles of a perturbed cavity
    agree with the FP64 formulation's, to 1e-9 for FP64 tables and 1e-5
    for the FP32 deviations of the mixed pipeline."""
    _say("[body] loads on table storage, cavity3d 12^3 nv6")
    case, cl, ocl, neigh, nn = setup("cavity3d", 12, 6, order=2)
    top = cl.z.max()
    surf = np.nonzero((cl.boundary == 1) & (cl.z == top))[0].astype(np.int32)
    centre = np.array([[0.3e-6, 0.45e-6, 0.5e-6]])
    area = [1e-15]
    out = {}
    for prec in (A.PREC_FP64, A.PREC_FP64_TABLE, A.PREC_MIXED):
        s = A.Solver(cl, precision=prec)
        s.step(case.dt)
        out[prec] = s.body_loads(surf, np.zeros(len(surf), np.int32), centre, area)[0]
        s.close()
    ref = out[A.PREC_FP64]
    _say(f"  FP64 {ref}\n  tab  {out[A.PREC_FP64_TABLE]}\n  mix  {out[A.PREC_MIXED]}")
    for prec, tol in ((A.PREC_FP64_TABLE, 1e-9), (A.PREC_MIXED, 1e-5)):
        for sl in (slice(0, 3), slice(3, 6)):
            den = np.abs(ref[sl]).max()
            assert den > 0
            assert np.abs(out[prec][sl] - ref[sl]).max() <= tol * den, (prec, sl)
