"""Restart files: the host library reads the reference's own restart files
(written by ckpt::save, Core/Checkpoint.hpp:78) field for field, and writes
files byte-identical to them."""
from __future__ import annotations

import json
import os

import numpy as np
import pytest

import alebgk4paper_amd as A

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def _ref_fields(name):
    d = os.path.join(GOLD, name)
    ld = lambda f: np.load(os.path.join(d, f))
    flags = np.column_stack([ld("boundary.npy"), ld("rec_flags.npy")]).astype(np.uint8)
    ints = ld("rec_ints.npy")
    ints = np.column_stack([ints[:, 0], ints[:, 1], ld("bodyid.npy"), ld("voxel.npy")]).astype(np.int32)
    return {"x": ld("x.npy"), "y": ld("y.npy"), "z": ld("z.npy"), "mom": ld("rec_mom.npy"),
            "mesh": ld("rec_mesh.npy"), "wall": ld("wall.npy"), "grid": ld("rec_grid.npy"),
            "ints": ints, "flags": flags, "g": ld("g.npy"), "bodies": ld("bodies.npy")}


@pytest.mark.parametrize("name", ["ckpt_cavity3d_n6_v5", "ckpt_cube_n8_v4"])
def test_reads_reference_restart_file(name):
    path = os.path.join(GOLD, name, "restart.ckpt")
    info = A.checkpoint_info(path)
    meta = json.load(open(os.path.join(GOLD, name, "meta.json")))
    assert (info["n"], info["gstride"], info["step"], info["t"], info["dim"]) == \
        (meta["N"], meta["gStride"], 7, 1.5e-10, 3)
    got = A.checkpoint_read(path)
    ref = _ref_fields(name)
    for k, v in ref.items():
        assert np.array_equal(got[k], v), k


@pytest.mark.parametrize("name", ["ckpt_cavity3d_n6_v5", "ckpt_cube_n8_v4"])
def test_writes_byte_identical_restart_file(name, tmp_path):
    ref = _ref_fields(name)
    out = tmp_path / "restart.ckpt"
    A.checkpoint_write(out, 7, 1.5e-10, ref)
    assert out.read_bytes() == open(os.path.join(GOLD, name, "restart.ckpt"), "rb").read()
    assert not (tmp_path / "restart.ckpt.tmp").exists()


def test_refuses_mismatched_or_truncated_files(tmp_path):
    path = os.path.join(GOLD, "ckpt_cavity3d_n6_v5", "restart.ckpt")
    info = A.checkpoint_info(path)
    with pytest.raises(A.ALEBGKError, match="unsupported"):
        A.checkpoint_read(path, n=info["n"] + 1)
    with pytest.raises(A.ALEBGKError):
        A.checkpoint_read(path, gstride=info["gstride"] * 2)
    with pytest.raises(A.ALEBGKError):
        A.checkpoint_read(path, nbodies=1)
    short = tmp_path / "short.ckpt"
    short.write_bytes(open(path, "rb").read()[:-8])
    with pytest.raises(A.ALEBGKError):
        A.checkpoint_read(short)
    bad = tmp_path / "bad.ckpt"
    bad.write_bytes(b"\0" * 64)
    with pytest.raises(A.ALEBGKError):
        A.checkpoint_info(bad)
