"""The discrete-Maxwellian collision of the oracle (CPU), pinned by what the
mathematics fixes: a grid Maxwellian is its own solution, the corrected
collision conserves the grid moments exactly, Newton converges
quadratically, and on a fine grid the solution tends to the continuous
Maxwellian's parameters."""
from __future__ import annotations

import numpy as np
import pytest

import oracle_ctypes as O
from helpers import setup

R = 208.0


def grid(Nv, vlo, dv):
    v = vlo + np.arange(Nv) * dv
    vz, vy, vx = np.meshgrid(v, v, v, indexing="ij")       # lin = iv + Nv (jv + Nv kv)
    return vx.ravel(), vy.ravel(), vz.ravel()


def scaled_moments(f, Nv, vlo, dv, cs):
    vx, vy, vz = grid(Nv, vlo, dv)
    xs = np.stack([np.ones_like(vx), vx / cs, vy / cs, vz / cs, (vx**2 + vy**2 + vz**2) / cs**2])
    return xs @ f * dv**3


@pytest.mark.parametrize("Nv", [6, 10, 16])
def test_grid_maxwellian_is_its_own_solution(Nv):
    vlo, vhi = -1303.0, 1303.0
    dv = (vhi - vlo) / (Nv - 1)
    cs = O.dm_scale(Nv, vlo, dv)
    rng = np.random.default_rng(Nv)
    for _ in range(5):
        T = 270.0 * (1 + 0.3 * rng.uniform(-1, 1))
        u = rng.uniform(-150, 150, 3)
        a_true = O.dm_guess(1.0, u, T, R, cs)
        a_true[1:4] += rng.uniform(-0.05, 0.05, 3)
        a_true[4] *= 1 + rng.uniform(-0.05, 0.05)
        vx, vy, vz = grid(Nv, vlo, dv)
        f = np.exp(a_true[0] + (a_true[1] * vx + a_true[2] * vy + a_true[3] * vz) / cs
                   + a_true[4] * (vx**2 + vy**2 + vz**2) / cs**2)
        m = scaled_moments(f, Nv, vlo, dv, cs)
        rho = m[0]; uu = m[1:4] * cs / rho
        Tm = (m[4] * cs**2 / rho - uu @ uu) / (3 * R)
        a, n, res = O.dm_solve(Nv, vlo, dv, cs, m, O.dm_guess(rho, uu, Tm, R, cs), 1e-13, 12)
        assert n <= 8 and res <= 1e-13
        np.testing.assert_allclose(a, a_true, rtol=0, atol=1e-9 * np.abs(a_true).max())


def test_newton_converges_quadratically():
    Nv, vlo, vhi = 10, -1303.0, 1303.0
    dv = (vhi - vlo) / (Nv - 1)
    cs = O.dm_scale(Nv, vlo, dv)
    vx, vy, vz = grid(Nv, vlo, dv)
    # a continuous Maxwellian sampled on the grid: its grid moments differ
    # from the continuous ones, so the guess is off by the quadrature error
    T, u = 300.0, np.array([80.0, -40.0, 20.0])
    f = np.exp(-((vx - u[0])**2 + (vy - u[1])**2 + (vz - u[2])**2) / (2 * R * T)) / (2 * np.pi * R * T)**1.5
    f *= 1 + 0.05 * np.sin(vx / 700.0)                        # and not a Maxwellian
    m = scaled_moments(f, Nv, vlo, dv, cs)
    rho = m[0]; uu = m[1:4] * cs / rho
    Tm = (m[4] * cs**2 / rho - uu @ uu) / (3 * R)
    a0 = O.dm_guess(rho, uu, Tm, R, cs)
    res = []
    a = a0.copy()
    for k in range(1, 6):
        a_k, n, r = O.dm_solve(Nv, vlo, dv, cs, m, a0, 0.0, k - 1)
        res.append(r)
    # residual after each pass: quadratic decrease until round-off
    assert res[0] > 1e-6
    for k in range(1, 3):
        assert res[k] < 1e3 * res[k - 1] ** 2 + 1e-14, res
    assert res[3] < 1e-14


def test_fine_grid_tends_to_continuous_parameters():
    """With 40 nodes per axis over +-5.5 thermal speeds the trapezoid rule is
    exact to ~1e-12 for a Gaussian: the discrete solution equals the
    continuous Maxwellian's parameters."""
    T = 270.0
    c = np.sqrt(2 * R * T)
    Nv, vlo = 40, -5.5 * c
    dv = -2 * vlo / (Nv - 1)
    cs = O.dm_scale(Nv, vlo, dv)
    u = np.array([30.0, -10.0, 5.0])
    vx, vy, vz = grid(Nv, vlo, dv)
    a_c = O.dm_guess(1.3, u, T, R, cs)
    f = np.exp(a_c[0] + (a_c[1] * vx + a_c[2] * vy + a_c[3] * vz) / cs
               + a_c[4] * (vx**2 + vy**2 + vz**2) / cs**2)
    # the continuous moments of that Maxwellian, in closed form
    rho, RT = 1.3, R * T
    mc = np.array([rho, rho * u[0] / cs, rho * u[1] / cs, rho * u[2] / cs,
                   rho * (u @ u + 3 * RT) / cs**2])
    m = scaled_moments(f, Nv, vlo, dv, cs)
    np.testing.assert_allclose(m, mc, rtol=1e-10, atol=1e-12)
    a, n, r = O.dm_solve(Nv, vlo, dv, cs, mc, O.dm_guess(rho, u, T, R, cs), 1e-13, 8)
    np.testing.assert_allclose(a, a_c, atol=1e-9)


def test_dm_collision_conserves_grid_moments():
    case, cl, ocl, neigh, nn = setup("cavity3d", 6, 8, order=2)
    Nv, vlo, dv = cl.cfg.nv, cl.cfg.vmin, (cl.cfg.vmax - cl.cfg.vmin) / (cl.cfg.nv - 1)
    cs = O.dm_scale(Nv, vlo, dv)
    g0 = cl.g
    g1, a, passes, res = O.collision_dm(ocl, g0, 1e-11, case.tau, case.R)
    inner = cl.boundary == 0
    m0 = np.array([scaled_moments(g0[i], Nv, vlo, dv, cs) for i in np.where(inner)[0]])
    m1 = np.array([scaled_moments(g1[i], Nv, vlo, dv, cs) for i in np.where(inner)[0]])
    np.testing.assert_allclose(m1, m0, rtol=0, atol=1e-12 * m0[:, 0].max())
    assert passes[inner].max() <= 4 and res[inner].max() <= 1e-12
    # the reference collision with the continuous Maxwellian does not
    mom = O.moments(ocl, g0, case.R, cl.mom)
    g2 = O.collision(ocl, g0, 1e-11, case.tau, case.R, mom)
    m2 = np.array([scaled_moments(g2[i], Nv, vlo, dv, cs) for i in np.where(inner)[0]])
    assert np.abs(m2 - m0)[:, 4].max() > 1e3 * max(np.abs(m1 - m0)[:, 4].max(), 1e-300)
    # boundary particles are untouched
    assert np.array_equal(g1[~inner], g0[~inner])


def test_warm_start_from_the_solution_takes_one_pass():
    case, cl, ocl, neigh, nn = setup("cavity3d", 6, 8, order=2)
    g1, a, passes, res = O.collision_dm(ocl, cl.g, 1e-11, case.tau, case.R)
    g2, a2, passes2, res2 = O.collision_dm(ocl, cl.g, 1e-11, case.tau, case.R, a_init=a)
    inner = cl.boundary == 0
    assert passes2[inner].max() == 1
    assert np.abs(g2 - g1).max() <= 1e-12 * np.abs(g1).max()
