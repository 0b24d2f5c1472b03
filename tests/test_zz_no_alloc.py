"""GPU: ab_body_loads_static, the body loads with the library's static
scratch (no device allocation), called through the C ABI, against
ab_body_loads on the same inputs: bit-identical results, on two calls in a
row (the second with fewer surface particles and bodies, so stale scratch
from the first call would show), for the uploaded wall velocities and for
given ones; sizes past the scratch are refused."""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

import alebgk4paper_amd as A
from conftest import load_fixture

pytestmark = pytest.mark.gpu

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)


def _call(fn, sol, idx, body, centres, areas, wu):
    nb = len(areas)
    out = np.zeros(6 * nb)
    idx = np.ascontiguousarray(idx, dtype=np.int32)
    body = np.ascontiguousarray(body, dtype=np.int32)
    c = np.ascontiguousarray(centres, dtype=np.float64).reshape(-1)
    a = np.ascontiguousarray(areas, dtype=np.float64)
    w = None if wu is None else np.ascontiguousarray(wu, dtype=np.float64).reshape(-1)
    rc = fn(sol._ctx, C.c_int(len(idx)), idx.ctypes.data_as(_ip), body.ctypes.data_as(_ip),
            C.c_int(nb), c.ctypes.data_as(_dp), a.ctypes.data_as(_dp),
            C.cast(None, _dp) if w is None else w.ctypes.data_as(_dp), out.ctypes.data_as(_dp))
    return rc, out.reshape(nb, 6)


def test_static_scratch_body_loads_match_allocating_entry_point():
    print("[no-alloc] ab_body_loads_static vs ab_body_loads, cube 12^3 nv4", flush=True)
    fx = load_fixture("bodysteps_cube_n12_v4")
    cs = A.configure("cavity3d-cube", nx=12, nv=4)
    cl = A.build_case(cs, body_scale=float(fx["params"][0]))
    rng = np.random.default_rng(17)
    cl.g = np.ascontiguousarray(fx["g0"] * (1.0 + 0.3 * rng.uniform(-1, 1, size=fx["g0"].shape)))
    sol = A.Solver(cl, precision=A.PREC_FP64)
    lib = A.hip_lib()
    surf = np.nonzero((cl.boundary == 1) & (cl.body_id >= 0))[0].astype(np.int32)
    b = cl.bodies[0]
    area = 6.0 * b.size * b.size / b.npts
    # first call: every surface particle split over three "bodies" (three
    # arms about different centres), given wall velocities
    body3 = (np.arange(len(surf)) % 3).astype(np.int32)
    cen3 = np.array([b.c[:], [5e-7, 5e-7, 5e-7], [6.5e-7, 6.8e-7, 4.9e-7]])
    wu = rng.normal(scale=4.0, size=(len(surf), 3))
    rs, s1 = _call(lib.ab_body_loads_static, sol, surf, body3, cen3, [area] * 3, wu)
    ra, a1 = _call(lib.ab_body_loads, sol, surf, body3, cen3, [area] * 3, wu)
    assert rs == 0 and ra == 0, lib.ab_last_error(sol._ctx)
    assert np.array_equal(s1, a1)
    assert np.all(np.abs(s1) > 0)
    # second call: a smaller input (a third of the surface, one body, the
    # uploaded wall velocities)
    sub = surf[::3]
    rs, s2 = _call(lib.ab_body_loads_static, sol, sub, np.zeros(len(sub)), [b.c[:]], [area], None)
    ra, a2 = _call(lib.ab_body_loads, sol, sub, np.zeros(len(sub)), [b.c[:]], [area], None)
    assert rs == 0 and ra == 0
    assert np.array_equal(s2, a2)
    assert not np.array_equal(s2[0], s1[0])
    # past the scratch: refused, and the context keeps working
    big = np.zeros(A.STATIC_SURF + 1, np.int32)
    rs, _ = _call(lib.ab_body_loads_static, sol, big, big, [b.c[:]], [area], None)
    assert rs == 1 and b"exceed" in lib.ab_last_error(sol._ctx)
    rs, s3 = _call(lib.ab_body_loads_static, sol, sub, np.zeros(len(sub)), [b.c[:]], [area], None)
    assert rs == 0 and np.array_equal(s3, a2)
