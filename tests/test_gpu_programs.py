"""GPU: the two programs. The command-line program alebgk_amd runs the
reference's cases on the device; the reference's own program, built from its
sources with Solver/CudaBackend.cu replaced by the binding
(alebgk4paper_amd/binding/HipBackend.cpp, recipe `make -C oracle hiplink`),
runs on this library as its GPU backend. Both write the reference's files:
for cavity3d 10^3, Nv 10, the 501 steps of tfinal 5e-9 (the last one the short tail step), the VTK file and the
timeseries row of step 500 are byte-identical between the two, and the
moments they hold agree with the library stepped directly (to the 7
significant digits the files carry)."""
from __future__ import annotations

import os
import subprocess

import numpy as np
import pytest

import alebgk4paper_amd as A
from conftest import ROOT

pytestmark = pytest.mark.gpu

CLI = os.path.join(ROOT, "alebgk4paper_amd", "bin", "alebgk_amd")
REFPROG = os.path.join(ROOT, "oracle", "_ref", "alebgk_hip")


def _say(msg):
    print(msg, flush=True)


def _vtk_field(path, name):
    lines = open(path).read().splitlines()
    n = int([ln for ln in lines if ln.startswith("POINTS")][0].split()[1])
    k = lines.index(f"SCALARS {name} double 1")
    return np.array([float(v) for v in lines[k + 2:k + 2 + n]])


def test_cli_runs_the_cavity_and_writes_the_reference_files(tmp_path):
    assert os.path.exists(CLI), "the command-line program is not built"
    _say("[cli] cavity3d 10^3 nv10, 500 FP64 steps")
    r = subprocess.run([CLI, "cavity3d", "--Nx", "10", "--Nv", "10", "--tfinal", "5e-9",
                        "--out", str(tmp_path)], stdout=subprocess.PIPE,
                       stderr=subprocess.STDOUT, text=True, timeout=300)
    _say(r.stdout[-600:])
    assert r.returncode == 0
    d = tmp_path / "cavity3d_Nx10_Nv10"
    assert (d / "output_000000.vtk").exists() and (d / "output_000500.vtk").exists()
    rows = open(d / "timeseries.csv").read().splitlines()
    assert rows[0] == "t,rho_avg,KE,T_avg,p_avg,mood_count" and len(rows) == 2
    assert "Done. Steps=501" in r.stdout
    # the library stepped directly: 500 steps of the same case
    cs = A.configure("cavity3d", nx=10, nv=10)
    cl = A.build_lattice(cs)
    s = A.Solver(cl)
    s.steps(cs.dt, 500)
    m = s.moments()
    rho = _vtk_field(d / "output_000500.vtk", "rho")
    T = _vtk_field(d / "output_000500.vtk", "T")
    assert np.abs(rho - m[:, 0]).max() <= 1e-6 * np.abs(m[:, 0]).max()
    assert np.abs(T - m[:, 4]).max() <= 1e-6 * m[:, 4].max()


def test_cli_body_case_runs(tmp_path):
    _say("[cli] cavity2d-circle 24^2 nv6, 20 steps with a free circle")
    r = subprocess.run([CLI, "cavity2d-circle", "--Nx", "24", "--Nv", "6", "--tfinal", "3e-10",
                        "--body-scale", "2.0", "--out", str(tmp_path)], stdout=subprocess.PIPE,
                       stderr=subprocess.STDOUT, text=True, timeout=300)
    _say(r.stdout[-400:])
    assert r.returncode == 0 and "Done. Steps=20" in r.stdout
    d = tmp_path / "cavity2d-circle_Nx24_Nv6"
    rows = open(d / "bodies.csv").read().splitlines()
    assert rows[0].startswith("t,body,cx,cy") and len(rows) == 2


@pytest.mark.skipif(not os.path.exists(REFPROG),
                    reason="the reference program on this backend was not built here")
def test_reference_program_runs_on_the_library(tmp_path):
    """The reference's main() and runSolver, unchanged, with this library as
    their GPU backend: the files of step 500 are those alebgk_amd writes."""
    _say("[refprog] reference program on the HIP backend, cavity3d 10^3 nv10")
    a, b = tmp_path / "ref", tmp_path / "cli"
    a.mkdir()
    r = subprocess.run([REFPROG, "cavity3d", "--Nx", "10", "--Nv", "10", "--tfinal", "5e-9"],
                       cwd=a, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True,
                       timeout=300)
    _say(r.stdout[-800:])
    assert r.returncode == 0 and "Done. Steps=501" in r.stdout
    r2 = subprocess.run([CLI, "cavity3d", "--Nx", "10", "--Nv", "10", "--tfinal", "5e-9",
                         "--out", str(b)], stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                        text=True, timeout=300)
    assert r2.returncode == 0
    ra = a / "output" / "cavity3d_Nx10_Nv10"
    rb = b / "cavity3d_Nx10_Nv10"
    for f in ("output_000500.vtk", "timeseries.csv"):
        assert open(ra / f, "rb").read() == open(rb / f, "rb").read(), f
    # the step lines the two print at step 500 agree too
    la = [ln for ln in r.stdout.splitlines() if ln.startswith("Step   500")]
    lb = [ln for ln in r2.stdout.splitlines() if ln.startswith("Step   500")]
    assert la and la == lb
