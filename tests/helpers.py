"""Shared inputs for the parity tests: a case built by the product's host
library, the same arrays wrapped for the oracle, and a smooth non-equilibrium
distribution to step from."""
from __future__ import annotations

import numpy as np

import alebgk4paper_amd as A
import oracle_ctypes as O


def geom_of(cfg, case):
    """oracle geometry matching an ab_config built by alh_build_lattice."""
    g, *_ = O.lattice(cfg.dim, case.nx, case.dom, Ny=case.ny, Nz=case.nz if cfg.dim == 3 else 1)
    return g


def local_maxwellian(cloud, case, amp=1.0):
    """A local Maxwellian whose density, velocity and temperature vary
    smoothly across the box (deterministic; the input the golden harness
    uses, restated in numpy)."""
    cfg = cloud.cfg
    L = 1e-6
    v = A.velocities(cfg)
    sx, sy, sz = cloud.x / L, cloud.y / L, cloud.z / L
    rho = case.rho0 * (1.0 + amp * 0.2 * np.sin(2 * np.pi * sx) * np.cos(np.pi * sy))
    ux = amp * 60.0 * np.sin(np.pi * sy)
    uy = -amp * 40.0 * np.sin(np.pi * sx)
    uz = amp * 30.0 * np.cos(np.pi * sz) if cfg.dim == 3 else np.zeros_like(sx)
    T = 270.0 * (1.0 + amp * 0.1 * np.sin(np.pi * sz + 0.3))
    R = cfg.R
    twoRT = 2.0 * R * T
    if cfg.dim == 3:
        e = ((v[None, :, 0] - ux[:, None]) ** 2 + (v[None, :, 1] - uy[:, None]) ** 2
             + (v[None, :, 2] - uz[:, None]) ** 2) / twoRT[:, None]
        g = (rho / (np.pi * twoRT) ** 1.5)[:, None] * np.exp(-e)
    else:
        e = ((v[None, :, 0] - ux[:, None]) ** 2 + (v[None, :, 1] - uy[:, None]) ** 2) / twoRT[:, None]
        g1 = (rho / (np.pi * twoRT))[:, None] * np.exp(-e)
        g = np.concatenate([g1, (0.5 * R * T)[:, None] * g1 * (1.0 + 0.05 * sx[:, None])], axis=1)
    mom = np.stack([rho, ux, uy, uz, T], axis=1)
    return np.ascontiguousarray(g), mom


def setup(name, nx, nv, order=2, precision=A.PREC_FP64, perturb=True, lid_scale=1.0):
    """Product cloud (+ perturbed g), oracle cloud on the same arrays, and the
    oracle's neighbour lists."""
    case = A.configure(name, nx=nx, nv=nv, mls_order=order)
    cl = A.build_lattice(case, precision=precision, mls_order=order)
    if lid_scale != 1.0:
        cl.wall[:, 3:6] *= lid_scale
    if perturb:
        cl.g, cl.mom = local_maxwellian(cl, case)
    geo = geom_of(cl.cfg, case)
    neigh, nn, _ = O.neighbours(geo, cl.x, cl.y, cl.z)
    ocl = O.Cloud(cl.cfg.dim, cl.cfg.nv, cl.cfg.vmin,
                  (cl.cfg.vmax - cl.cfg.vmin) / (cl.cfg.nv - 1), cl.x, cl.y, cl.z,
                  cl.boundary, cl.wall, neigh, nn)
    return case, cl, ocl, neigh, nn


def rel_err(a, b, floor=0.0):
    a, b = np.asarray(a), np.asarray(b)
    den = np.maximum(np.abs(b).max(), floor)
    return float(np.abs(a - b).max() / den) if den > 0 else float(np.abs(a - b).max())


def moment_errors(m, ref, interior):
    """Relative errors of rho, |u| and T as north-star defines them: max over
    particles of |drho|/rho, |du|/max|u|, |dT|/T."""
    d = {}
    d["rho"] = float(np.max(np.abs(m[interior, 0] - ref[interior, 0]) / ref[interior, 0]))
    umax = np.max(np.abs(ref[interior, 1:4]))
    d["u"] = float(np.max(np.abs(m[interior, 1:4] - ref[interior, 1:4])) / max(umax, 1e-300))
    d["T"] = float(np.max(np.abs(m[interior, 4] - ref[interior, 4]) / ref[interior, 4]))
    return d
