"""CPU tests of the host-side run outputs, the stability guard, the
command-line program and the drop-in binding.

  writers   the VTK file and the timeseries rows byte-identical to what the
            reference's writeVTK / writeTimeseries wrote for the same cloud
            (fixtures writers_*: the reference's Core/Output.hpp run by the
            harness); computeDiagnostics' means identical
  guard     each bound of stability::Guard trips, the KE reference is the
            first positive value, the environment overrides and the switch-off
            are honoured
  CLI       the reference's options and refusals; without a GPU it fails
            loudly at the device context (no CPU fallback)
  binding   the reference's own sources, with Solver/CudaBackend.cu replaced
            by alebgk4paper_amd/binding/HipBackend.cpp, compile and link
            against libalebgk_hip.so (needs the reference sources: skipped
            where they are absent)
"""
from __future__ import annotations

import os
import subprocess
import sys

import numpy as np
import pytest

import alebgk4paper_amd as A
from conftest import ROOT, load_fixture

CLI = os.path.join(ROOT, "alebgk4paper_amd", "bin", "alebgk_amd")
REF = "/root/reference"


@pytest.mark.parametrize("name", ["writers_cavity3d_n6_v4", "writers_cavity2d_n8_v4"])
def test_writers_match_reference_bytes(name, tmp_path):
    fx = load_fixture(name)
    f = fx["fields"]
    x, y, z, rho, T, p, ux, uy, uz = [np.ascontiguousarray(f[:, k]) for k in range(9)]
    b, m = fx["boundary"], fx["mood"]
    d = A.diagnostics(b, rho, ux, uy, uz, T, p, m)
    ref = fx["diag"]
    assert (d.rho_avg, d.KE, d.T_avg, d.p_avg, d.mood_count, d.n) == tuple(
        [ref[0], ref[1], ref[2], ref[3], int(ref[4]), int(ref[5])])
    A.write_vtk(tmp_path, 7, 7e-11, x, y, z, rho, T, p, fx.meta["Nv"], m, b, ux, uy, uz)
    got = open(tmp_path / "output_000007.vtk", "rb").read()
    want = open(os.path.join(fx.dir, "output_000007.vtk"), "rb").read()
    assert got == want
    ts = tmp_path / "timeseries.csv"
    A.write_timeseries(ts, True, 1e-10, d)
    A.write_timeseries(ts, False, 2e-10, d)
    assert open(ts, "rb").read() == open(os.path.join(fx.dir, "timeseries.csv"), "rb").read()


def test_step_line_and_body_log(tmp_path):
    d = A.Diag(rho_avg=1.0, KE=2.5e-3, T_avg=270.1234, p_avg=0.0, mood_count=7, n=10)
    line = A.format_step(500, 5e-9, 1e-11, d)
    assert line == "Step   500|t=5.000e-09|dt=1.00e-11|rho= 1.0000|T=270.1234|KE=2.500e-03|MOOD=  7"
    cs = A.configure("cavity3d-cube", nx=12, nv=4)
    bodies = A.make_bodies(cs)
    bodies[0].vel[:] = [1.5, -2.0, 0.25]
    bodies[0].om[:] = [10.0, 20.0, 30.0]
    path = tmp_path / "bodies.csv"
    A.write_bodies(path, True, 1e-9, 3, bodies)
    A.write_bodies(path, False, 2e-9, 3, bodies)
    rows = open(path).read().splitlines()
    assert rows[0] == "t,body,cx,cy,velx,vely,angVel,Fx,Fy,torque,cz,velz,omx,omy,Fz,tqx,tqy"
    assert len(rows) == 3
    r = rows[1].split(",")
    assert r[0] == "1.000000e-09" and r[1] == "0"
    assert r[6] == "3.000000e+01"          # angVel is omega_z in 3-D
    assert r[12] == "1.000000e+01" and r[13] == "2.000000e+01"


def _diag(rho=1.0, T=270.0, KE=1.0):
    return A.Diag(rho_avg=rho, KE=KE, T_avg=T, p_avg=0.0, mood_count=0, n=1)


def test_guard_bounds():
    g = A.Guard(270.0, 4.4)
    assert g.diverged(1, 0.0, _diag(rho=4.4, KE=0.0)) == (False, "")
    # the KE reference is the first positive value
    assert not g.diverged(2, 0.0, _diag(rho=4.4, KE=2.0))[0]
    assert not g.diverged(3, 0.0, _diag(rho=4.4, KE=199.0))[0]
    bad, why = g.diverged(4, 0.0, _diag(rho=4.4, KE=201.0))
    assert bad and "kinetic energy" in why
    for kw, word in ((dict(T=0.59 * 270), "below 162 K"), (dict(T=1.31 * 270), "above 351 K"),
                     (dict(rho=0.19 * 4.4), "density below"), (dict(rho=5.01 * 4.4), "density above"),
                     (dict(T=float("nan")), "NaN")):
        bad, why = A.Guard(270.0, 4.4).diverged(1, 0.0, _diag(**{"rho": 4.4, **kw}))
        assert bad and word in why, (kw, why)
    assert not A.Guard(270.0, 4.4).diverged(1, 0.0, _diag(rho=4.4, T=1.29 * 270))[0]


def test_guard_environment_overrides():
    code = ("import alebgk4paper_amd as A\n"
            "g = A.Guard(270.0, 1.0)\n"
            "d = A.Diag(rho_avg=1.0, KE=1.0, T_avg=1.4 * 270, p_avg=0, mood_count=0, n=1)\n"
            "print(int(g.diverged(1, 0.0, d)[0]))\n")
    env = dict(os.environ, PYTHONPATH=ROOT)
    run = lambda extra: subprocess.run([sys.executable, "-c", code], env={**env, **extra},
                                       stdout=subprocess.PIPE, text=True, check=True).stdout.strip()
    assert run({}) == "1"
    assert run({"ALEBGK_GUARD_TMAX": "1.5"}) == "0"
    assert run({"ALEBGK_NO_GUARD": "1"}) == "0"


def _cli(*args, cwd=None):
    return subprocess.run([CLI, *args], stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                          text=True, cwd=cwd, timeout=120)


def test_cli_options_and_refusals(tmp_path):
    assert os.path.exists(CLI), "the command-line program is not built"
    r = _cli()
    assert r.returncode == 1 and "Usage: alebgk_amd <case>" in r.stdout
    assert "cavity3d-cube" in r.stdout and "--body-scale" in r.stdout
    r = _cli("cavity4d")
    assert r.returncode == 1 and "Unknown case 'cavity4d'" in r.stdout
    r = _cli("cavity3d", "--slab")
    assert r.returncode == 1 and "MPI" in r.stdout
    r = _cli("cavity3d", "--precision", "fp16")
    assert r.returncode == 1 and "needs a valid value" in r.stdout
    r = _cli("cavity3d", "--Nx")
    assert r.returncode == 1
    r = _cli("cavity3d", "--frobnicate")
    assert r.returncode == 1 and "Unknown option" in r.stdout


def test_cli_fails_loudly_without_gpu(tmp_path):
    try:
        import torch
        if torch.cuda.is_available():
            pytest.skip("a GPU is present")
    except ImportError:
        pass
    r = _cli("6", "--Nx", "10", "--Nv", "4", "--out", str(tmp_path))
    assert r.returncode == 1 and "ab_create" in r.stdout
    assert "Case 6: cavity3d-cube" in r.stdout and "free to translate" in r.stdout
    # the output directory of the reference's naming exists
    assert (tmp_path / "cavity3d-cube_Nx10_Nv4").is_dir()


@pytest.mark.skipif(not os.path.isdir(os.path.join(REF, "Solver")),
                    reason="the reference sources are not present")
def test_reference_links_against_the_hip_library():
    r = subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "hiplink"],
                       stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=900)
    assert r.returncode == 0, r.stdout[-2000:]
    exe = os.path.join(ROOT, "oracle", "_ref", "alebgk_hip")
    ldd = subprocess.run(["ldd", exe], stdout=subprocess.PIPE, text=True).stdout
    assert "libalebgk_hip.so" in ldd
    nm = subprocess.run(["nm", "-C", exe], stdout=subprocess.PIPE, text=True).stdout
    defined = {ln.split(" T ", 1)[1].split("(")[0] for ln in nm.splitlines() if " T " in ln}
    undefined = {ln.split(" U ", 1)[1].strip() for ln in nm.splitlines() if " U " in ln}
    for f in ("cudaUploadCloud", "cudaStep", "cudaDownloadMoments", "cudaDownloadFull",
              "cudaSyncBodies", "cudaLimiterCounts", "cudaMlsCounts", "cudaMlsOvershoot",
              "cudaMlsOvershootNow", "cudaTotalMass", "cudaMassBudget", "cudaWallMomentum"):
        assert f in defined, f
    for f in ("ab_create", "ab_upload", "ab_step", "ab_download_moments", "ab_set_wall_state"):
        assert f in undefined, f
    run = subprocess.run([exe], stdout=subprocess.PIPE, text=True)
    assert run.returncode == 1 and "Usage: alebgk <case>" in run.stdout
