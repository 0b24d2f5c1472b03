"""Shared pytest configuration: the gpu marker and golden-fixture loaders."""
from __future__ import annotations

import json
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
TESTS = os.path.join(ROOT, "tests")
if TESTS not in sys.path:
    sys.path.insert(0, TESTS)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


class Fixture:
    """One golden fixture directory written by tests/golden/make_golden.sh."""

    def __init__(self, name):
        self.name = name
        self.dir = os.path.join(GOLDEN, name)
        with open(os.path.join(self.dir, "meta.json")) as f:
            self.meta = json.load(f)

    def __getitem__(self, key):
        return np.load(os.path.join(self.dir, key + ".npy"))

    def has(self, key):
        return os.path.exists(os.path.join(self.dir, key + ".npy"))

    def dom(self):
        m = self.meta
        return (m["xleft"], m["xright"], m["ytop"], m["ybottom"], m["zfront"], m["zback"])

    def neigh_padded(self, max_neigh=128):
        nn = self["nneigh"]
        flat = self["neigh_flat"]
        out = np.full((len(nn), max_neigh), -1, np.int32)
        off = np.concatenate([[0], np.cumsum(nn)])
        for i in range(len(nn)):
            out[i, : nn[i]] = flat[off[i]: off[i + 1]]
        return out


def load_fixture(name):
    return Fixture(name)


@pytest.fixture(scope="session")
def c3d():
    return Fixture("cavity3d_n6_v5")


@pytest.fixture(scope="session")
def c3d9():
    return Fixture("cavity3d_n9_v3")


@pytest.fixture(scope="session")
def c2d():
    return Fixture("cavity2d_n10_v8")


@pytest.fixture(scope="session")
def cube():
    return Fixture("cube_n12_v5")
