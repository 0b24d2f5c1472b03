"""GPU: a restart file written by the reference restores onto the device,
steps like the oracle from the same state, and saves back byte-identically."""
from __future__ import annotations

import os

import numpy as np
import pytest

import alebgk4paper_amd as A
import oracle_ctypes as O
from helpers import geom_of, rel_err

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(__file__), "golden")


def _say(msg):
    print(msg, flush=True)


def test_restore_steps_like_oracle_and_saves_identically(tmp_path):
    _say("[ckpt] restore cavity3d 6^3 nv5 from the reference's restart file")
    path = os.path.join(GOLD, "ckpt_cavity3d_n6_v5", "restart.ckpt")
    case = A.configure("cavity3d", nx=6, nv=5, mls_order=2)
    cl = A.cloud_from_checkpoint(path, case, precision=A.PREC_FP64, mls_order=2)
    sol = A.Solver(cl, precision=A.PREC_FP64)
    # unchanged state saves back byte for byte (voxels from the device rebuild)
    out = tmp_path / "again.ckpt"
    sol.save_checkpoint(out, cl.record["step"], cl.record["t"])
    assert out.read_bytes() == open(path, "rb").read()
    # two steps from the restored state, device vs oracle
    nb, nn = sol.neighbours()
    ocl = O.Cloud(3, cl.cfg.nv, cl.cfg.vmin, (cl.cfg.vmax - cl.cfg.vmin) / (cl.cfg.nv - 1),
                  cl.x, cl.y, cl.z, cl.boundary, cl.wall, nb, nn, mesh=cl.mesh)
    st = O.State(ocl, cl.g, cl.mom)
    dt = 0.5 * (1e-6 / 5) / 1303.0
    for _ in range(2):
        sol.step(dt)
        O.step(ocl, st, dt, case.tau, case.R, cl.cfg.radius, 2)
    assert rel_err(sol.g(), st.g) < 1e-10
