"""ctypes binding of oracle/liboracle.so -- the CPU FP64 checker.

Test infrastructure only: tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg use it to check the HIP library; the product never does.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(ROOT, "oracle", "liboracle.so")
MAX_NEIGH = 128

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)
_up = C.POINTER(C.c_ubyte)


class OrcGeom(C.Structure):
    _fields_ = [("dim", C.c_int), ("Nx", C.c_int), ("Ny", C.c_int), ("Nz", C.c_int),
                ("xleft", C.c_double), ("xright", C.c_double), ("ybottom", C.c_double),
                ("ytop", C.c_double), ("zfront", C.c_double), ("zback", C.c_double),
                ("dx", C.c_double), ("dy", C.c_double), ("dz", C.c_double),
                ("radius", C.c_double), ("minDist", C.c_double),
                ("hBox", C.c_double), ("vBox", C.c_double), ("wBox", C.c_double),
                ("nbhBox", C.c_int), ("nbvBox", C.c_int), ("nbwBox", C.c_int),
                ("nvox", C.c_int)]


class OrcCloud(C.Structure):
    _fields_ = [("dim", C.c_int), ("N", C.c_int), ("Nv", C.c_int), ("gstride", C.c_int),
                ("vlo", C.c_double), ("dv", C.c_double),
                ("x", _dp), ("y", _dp), ("z", _dp), ("boundary", _up),
                ("wnx", _dp), ("wny", _dp), ("wnz", _dp), ("wUx", _dp), ("wUy", _dp),
                ("wUz", _dp), ("wT", _dp), ("mvx", _dp), ("mvy", _dp), ("mvz", _dp),
                ("neigh", _ip), ("nneigh", _ip)]


class OrcState(C.Structure):
    _fields_ = [("g", _dp), ("gt", _dp), ("rho", _dp), ("ux", _dp), ("uy", _dp),
                ("uz", _dp), ("T", _dp), ("limflag", _dp), ("mlsflag", _ip),
                ("mlsovr", _dp)]


def _ptr(a, t):
    if a is None:
        return C.cast(None, t)
    assert a.flags["C_CONTIGUOUS"], "arrays passed to the oracle must be contiguous"
    return a.ctypes.data_as(t)


_lib = None


def lib():
    """Load liboracle.so, building it with make if it is missing."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "oracle"],
                           check=True)
        L = C.CDLL(LIB_PATH)
        L.orc_lattice.restype = C.c_int
        L.orc_neighbours.restype = C.c_int
        L.orc_sum_mass.restype = C.c_double
        L.orc_step.restype = C.c_double
        L.orc_step_dm.restype = C.c_double
        L.orc_step_mood.restype = C.c_double
        L.orc_dm_scale.restype = C.c_double
        L.orc_dm_scale.argtypes = [C.c_int, C.c_double, C.c_double]
        L.orc_maxwellian3d.restype = C.c_double
        L.orc_maxwellian3d.argtypes = [C.c_double] * 9
        L.orc_maxwellian_g1.restype = C.c_double
        L.orc_maxwellian_g1.argtypes = [C.c_double] * 7
        _lib = L
    return _lib


class Cloud:
    """A cloud in structure-of-arrays form, owning its numpy buffers."""

    def __init__(self, dim, Nv, vlo, dv, x, y, z, boundary, wall, neigh, nneigh,
                 mesh=None):
        self.dim, self.Nv, self.vlo, self.dv = int(dim), int(Nv), float(vlo), float(dv)
        self.N = int(len(x))
        self.gstride = Nv ** 3 if dim == 3 else 2 * Nv * Nv
        f8 = lambda a: np.ascontiguousarray(a, dtype=np.float64)
        self.x, self.y, self.z = f8(x), f8(y), f8(z)
        self.boundary = np.ascontiguousarray(boundary, dtype=np.uint8)
        wall = f8(wall).reshape(self.N, 7)
        self.wall = wall
        self.wn = [f8(wall[:, k]) for k in range(7)]
        self.mesh = None if mesh is None else [f8(mesh[:, k]) for k in range(3)]
        self.neigh = np.ascontiguousarray(neigh, dtype=np.int32).reshape(self.N, MAX_NEIGH)
        self.nneigh = np.ascontiguousarray(nneigh, dtype=np.int32)
        self._s = OrcCloud()
        s = self._s
        s.dim, s.N, s.Nv, s.gstride, s.vlo, s.dv = (self.dim, self.N, self.Nv, self.gstride,
                                                     self.vlo, self.dv)
        s.x, s.y, s.z = _ptr(self.x, _dp), _ptr(self.y, _dp), _ptr(self.z, _dp)
        s.boundary = _ptr(self.boundary, _up)
        (s.wnx, s.wny, s.wnz, s.wUx, s.wUy, s.wUz, s.wT) = [_ptr(a, _dp) for a in self.wn]
        if self.mesh is None:
            s.mvx = s.mvy = s.mvz = C.cast(None, _dp)
        else:
            s.mvx, s.mvy, s.mvz = [_ptr(a, _dp) for a in self.mesh]
        s.neigh, s.nneigh = _ptr(self.neigh, _ip), _ptr(self.nneigh, _ip)

    @property
    def ref(self):
        return C.byref(self._s)

    def velocities(self):
        """Velocity of every node, shape (gstride_nodes, 3), in lin order."""
        Nv, v = self.Nv, self.vlo + self.dv * np.arange(self.Nv)
        if self.dim == 3:
            kv, jv, iv = np.meshgrid(np.arange(Nv), np.arange(Nv), np.arange(Nv), indexing="ij")
            return np.stack([v[iv].ravel(), v[jv].ravel(), v[kv].ravel()], axis=1)
        jv, iv = np.meshgrid(np.arange(Nv), np.arange(Nv), indexing="ij")
        return np.stack([v[iv].ravel(), v[jv].ravel(), np.zeros(Nv * Nv)], axis=1)


# ---- setup -----------------------------------------------------------------

def lattice(dim, Nx, dom, Ny=None, Nz=None):
    """orc_lattice: returns (geom, x, y, z, boundary, voxel)."""
    g = OrcGeom()
    g.dim, g.Nx = dim, Nx
    g.Ny = Ny if Ny else Nx
    g.Nz = (Nz if Nz else Nx) if dim == 3 else 1
    (g.xleft, g.xright, g.ytop, g.ybottom, g.zfront, g.zback) = dom
    N = g.Nx * g.Ny * g.Nz
    x, y, z = np.zeros(N), np.zeros(N), np.zeros(N)
    b = np.zeros(N, np.uint8)
    vox = np.zeros(N, np.int32)
    n = lib().orc_lattice(C.byref(g), _ptr(x, _dp), _ptr(y, _dp), _ptr(z, _dp),
                          _ptr(b, _up), _ptr(vox, _ip))
    assert n == N
    return g, x, y, z, b, vox


def neighbours(geom, x, y, z):
    N = len(x)
    vox = np.zeros(N, np.int32)
    neigh = np.zeros((N, MAX_NEIGH), np.int32)
    nn = np.zeros(N, np.int32)
    rc = lib().orc_neighbours(C.byref(geom), C.c_int(N), _ptr(x, _dp), _ptr(y, _dp),
                              _ptr(z, _dp), _ptr(vox, _ip), _ptr(neigh, _ip), _ptr(nn, _ip))
    assert rc == 0
    return neigh, nn, vox


def wall_state(geom, x, y, z, boundary, bodyid, Tw, Uface):
    N = len(x)
    out = [np.zeros(N) for _ in range(7)]
    U = np.ascontiguousarray(Uface, dtype=np.float64).reshape(18)
    bid = None if bodyid is None else np.ascontiguousarray(bodyid, dtype=np.int32)
    lib().orc_wall_state(C.byref(geom), C.c_int(N), _ptr(x, _dp), _ptr(y, _dp), _ptr(z, _dp),
                         _ptr(boundary, _up), _ptr(bid, _ip), C.c_double(Tw), _ptr(U, _dp),
                         *[_ptr(a, _dp) for a in out])
    return np.stack(out, axis=1)


def lid_parabolic(geom, x, y, boundary, bodyid, peak, wUx):
    bid = None if bodyid is None else np.ascontiguousarray(bodyid, dtype=np.int32)
    lib().orc_lid_parabolic(C.byref(geom), C.c_int(len(x)), _ptr(x, _dp), _ptr(y, _dp),
                            _ptr(boundary, _up), _ptr(bid, _ip), C.c_double(peak),
                            _ptr(wUx, _dp))


def initial_state(dim, N, Nv, vlo, dv, rho0, T0, R):
    gs = Nv ** 3 if dim == 3 else 2 * Nv * Nv
    g = np.zeros((N, gs))
    lib().orc_initial_state(C.c_int(dim), C.c_int(N), C.c_int(Nv), C.c_double(vlo),
                            C.c_double(dv), C.c_double(rho0), C.c_double(T0), C.c_double(R),
                            _ptr(g, _dp))
    return g


# ---- kernels ------------------------------------------------------------------

def transport(cloud, g, dt, h, order, tol=1e-12, clamp=0, flags=False):
    g = np.ascontiguousarray(g, dtype=np.float64)
    gt = np.zeros_like(g)
    if cloud.dim == 3:
        fl = np.zeros(cloud.N, np.int32) if flags else None
        ov = np.zeros(cloud.N) if flags else None
        lib().orc_transport3d(cloud.ref, _ptr(g, _dp), _ptr(gt, _dp), C.c_double(dt),
                              C.c_double(h), C.c_int(order), C.c_double(tol), C.c_int(clamp),
                              _ptr(fl, _ip), _ptr(ov, _dp))
        return (gt, fl, ov) if flags else gt
    lib().orc_transport2d(cloud.ref, _ptr(g, _dp), _ptr(gt, _dp), C.c_double(dt),
                          C.c_double(h), C.c_int(order))
    return gt


def positivity(cloud, g, gt):
    gt = np.ascontiguousarray(gt, dtype=np.float64).copy()
    lim = np.zeros(cloud.N)
    if cloud.dim == 3:
        g = np.ascontiguousarray(g, dtype=np.float64)
        lib().orc_positivity3d(cloud.ref, _ptr(g, _dp), _ptr(gt, _dp), _ptr(lim, _dp))
    else:
        lib().orc_positivity2d(cloud.ref, _ptr(gt, _dp), _ptr(lim, _dp))
    return gt, lim


def nad(cloud, gt, rho):
    gt = np.ascontiguousarray(gt, dtype=np.float64).copy()
    rho = np.ascontiguousarray(rho, dtype=np.float64)
    lim = np.zeros(cloud.N)
    lib().orc_nad3d(cloud.ref, _ptr(gt, _dp), _ptr(rho, _dp), _ptr(lim, _dp))
    return gt, lim


def sum_mass(cloud, g):
    g = np.ascontiguousarray(g, dtype=np.float64)
    return lib().orc_sum_mass(cloud.ref, _ptr(g, _dp))


def wallbc(cloud, g, R, mom):
    g = np.ascontiguousarray(g, dtype=np.float64).copy()
    m = [np.ascontiguousarray(mom[:, k]).copy() for k in range(5)]
    if cloud.dim == 3:
        lib().orc_wallbc3d(cloud.ref, _ptr(g, _dp), C.c_double(R), *[_ptr(a, _dp) for a in m])
    else:
        lib().orc_wallbc2d(cloud.ref, _ptr(g, _dp), C.c_double(R), _ptr(m[0], _dp),
                           _ptr(m[1], _dp), _ptr(m[2], _dp), _ptr(m[4], _dp))
    return g, np.stack(m, axis=1)


def moments(cloud, g, R, mom):
    g = np.ascontiguousarray(g, dtype=np.float64)
    m = [np.ascontiguousarray(mom[:, k]).copy() for k in range(5)]
    if cloud.dim == 3:
        lib().orc_moments3d(cloud.ref, _ptr(g, _dp), C.c_double(R), *[_ptr(a, _dp) for a in m])
    else:
        lib().orc_moments2d(cloud.ref, _ptr(g, _dp), C.c_double(R), _ptr(m[0], _dp),
                            _ptr(m[1], _dp), _ptr(m[2], _dp), _ptr(m[4], _dp))
    return np.stack(m, axis=1)


def collision(cloud, g, dt, tau, R, mom):
    g = np.ascontiguousarray(g, dtype=np.float64).copy()
    m = [np.ascontiguousarray(mom[:, k]) for k in range(5)]
    if cloud.dim == 3:
        lib().orc_collision3d(cloud.ref, _ptr(g, _dp), C.c_double(dt), C.c_double(tau),
                              C.c_double(R), *[_ptr(a, _dp) for a in m])
    else:
        lib().orc_collision2d(cloud.ref, _ptr(g, _dp), C.c_double(dt), C.c_double(tau),
                              C.c_double(R), _ptr(m[0], _dp), _ptr(m[1], _dp),
                              _ptr(m[2], _dp), _ptr(m[4], _dp))
    return g


class State:
    """Mutable step state for orc_step."""

    def __init__(self, cloud, g, mom):
        self.g = np.ascontiguousarray(g, dtype=np.float64).copy()
        self.gt = np.zeros_like(self.g)
        self.m = [np.ascontiguousarray(mom[:, k], dtype=np.float64).copy() for k in range(5)]
        self.limflag = np.zeros(cloud.N)
        self.mlsflag = np.zeros(cloud.N, np.int32)
        self.mlsovr = np.zeros(cloud.N)
        s = OrcState()
        s.g, s.gt = _ptr(self.g, _dp), _ptr(self.gt, _dp)
        s.rho, s.ux, s.uy, s.uz, s.T = [_ptr(a, _dp) for a in self.m]
        s.limflag, s.mlsflag, s.mlsovr = (_ptr(self.limflag, _dp), _ptr(self.mlsflag, _ip),
                                          _ptr(self.mlsovr, _dp))
        self._s = s

    @property
    def mom(self):
        return np.stack(self.m, axis=1)


def step(cloud, state, dt, tau, R, h, order, tol=1e-12, nofix=0):
    return lib().orc_step(cloud.ref, C.byref(state._s), C.c_double(dt), C.c_double(tau),
                          C.c_double(R), C.c_double(h), C.c_int(order), C.c_double(tol),
                          C.c_int(nofix))


def step_dm(cloud, state, dt, tau, R, h, order, tol=1e-12, nofix=0, dm_tol=1e-12, dm_maxit=8):
    return lib().orc_step_dm(cloud.ref, C.byref(state._s), C.c_double(dt), C.c_double(tau),
                             C.c_double(R), C.c_double(h), C.c_int(order), C.c_double(tol),
                             C.c_int(nofix), C.c_double(dm_tol), C.c_int(dm_maxit))


def step_mood(cloud, state, dt, tau, R, h, order, tol=1e-12):
    return lib().orc_step_mood(cloud.ref, C.byref(state._s), C.c_double(dt), C.c_double(tau),
                               C.c_double(R), C.c_double(h), C.c_int(order), C.c_double(tol))


def dm_scale(Nv, vlo, dv):
    return lib().orc_dm_scale(Nv, vlo, dv)


def dm_guess(rho, u, T, R, cs):
    a = np.zeros(5)
    lib().orc_dm_guess(C.c_double(rho), C.c_double(u[0]), C.c_double(u[1]), C.c_double(u[2]),
                       C.c_double(T), C.c_double(R), C.c_double(cs), _ptr(a, _dp))
    return a


def dm_solve(Nv, vlo, dv, cs, m, a, tol=1e-12, maxit=8):
    """Returns (a, passes, residual)."""
    m = np.ascontiguousarray(m, dtype=np.float64)
    a = np.ascontiguousarray(a, dtype=np.float64).copy()
    r = C.c_double(0.0)
    n = lib().orc_dm_solve(C.c_int(Nv), C.c_double(vlo), C.c_double(dv), C.c_double(cs),
                           _ptr(m, _dp), _ptr(a, _dp), C.c_double(tol), C.c_int(maxit),
                           C.byref(r))
    return a, n, r.value


def collision_dm(cloud, g, dt, tau, R, tol=1e-12, maxit=8, a_init=None):
    """Returns (g, alpha N x 5, passes N, residual N)."""
    g = np.ascontiguousarray(g, dtype=np.float64).copy()
    N = cloud.N
    a_out = np.zeros((N, 5)); passes = np.zeros(N, np.int32); res = np.zeros(N)
    ainit = None if a_init is None else np.ascontiguousarray(a_init, dtype=np.float64)
    ai = None if ainit is None else _ptr(ainit, _dp)
    lib().orc_collision3d_dm(cloud.ref, _ptr(g, _dp), C.c_double(dt), C.c_double(tau),
                             C.c_double(R), C.c_double(tol), C.c_int(maxit), ai,
                             _ptr(a_out, _dp), _ptr(passes, _ip), _ptr(res, _dp))
    return g, a_out, passes, res


def body_loads(cloud, g, bodyid, body, centre, area):
    g = np.ascontiguousarray(g, dtype=np.float64)
    bid = np.ascontiguousarray(bodyid, dtype=np.int32)
    out = np.zeros(6)
    lib().orc_body_loads(cloud.ref, _ptr(g, _dp), _ptr(bid, _ip), C.c_int(body),
                         C.c_double(centre[0]), C.c_double(centre[1]), C.c_double(centre[2]),
                         C.c_double(area), _ptr(out, _dp))
    return out


# ---- cloud management around moving bodies ----------------------------------

class OrcBody(C.Structure):
    _fields_ = [("shape", C.c_int), ("c", C.c_double * 3), ("size", C.c_double),
                ("o", C.c_double * 2), ("rm", C.c_double * 9)]


def body_from_vector(b):
    """An OrcBody from (shape, cx, cy, cz, size, ox, oy, rm[9])."""
    o = OrcBody()
    o.shape = int(b[0])
    o.c[:] = [float(v) for v in b[1:4]]
    o.size = float(b[4])
    o.o[:] = [float(v) for v in b[5:7]]
    o.rm[:] = [float(v) for v in b[7:16]]
    return o


def geom_from_meta(m):
    g = OrcGeom()
    g.dim, g.Nx, g.Ny, g.Nz = m["dim"], m["Nx"], m["Ny"], m["Nz"]
    for k in ("xleft", "xright", "ybottom", "ytop", "zfront", "zback", "dx", "dy", "dz",
              "radius", "minDist"):
        setattr(g, k, m[k])
    g.hBox, g.vBox, g.wBox = m["hBox"], m["vBox"], m["wBox"]
    g.nbhBox, g.nbvBox, g.nbwBox, g.nvox = m["nbhBox"], m["nbvBox"], m["nbwBox"], m["nvox"]
    return g


def cloud_manage(m, x, y, z, boundary, bodyid, g, mom, bodies):
    """The cloud management of the moving-body block as the reference
    sequences it (Solver/Solver.cpp:379-398) on the oracle: remove,
    neighbours, merge, refill. Returns the new cloud (x, y, z, boundary,
    bodyid, g, mom), src (index before, -1 refilled) and the counts."""
    L = lib()
    d = int(m["dim"])
    f8 = lambda a: np.ascontiguousarray(a, dtype=np.float64).copy()
    x, y, z, g, mom = f8(x), f8(y), f8(z), f8(g), f8(mom)
    b = np.ascontiguousarray(boundary, dtype=np.uint8).copy()
    bid = np.ascontiguousarray(bodyid, dtype=np.int32).copy()
    src = np.arange(len(x), dtype=np.int32)
    nb = len(bodies)
    barr = (OrcBody * max(nb, 1))(*bodies)
    dx, geom = float(m["dx"]), geom_from_meta(m)
    keep = np.zeros(len(x), np.uint8)
    removed = L.orc_cloud_remove(d, len(x), _ptr(x, _dp), _ptr(y, _dp), _ptr(z, _dp),
                                 _ptr(b, _up), _ptr(bid, _ip), nb, barr, C.c_double(dx),
                                 _ptr(keep, _up))

    def take(sel):
        nonlocal x, y, z, b, bid, g, mom, src
        x, y, z = [np.ascontiguousarray(a[sel]) for a in (x, y, z)]
        b, bid, src = [np.ascontiguousarray(a[sel]) for a in (b, bid, src)]
        g, mom = np.ascontiguousarray(g[sel]), np.ascontiguousarray(mom[sel])

    take(keep == 1)
    neigh, nn, _ = neighbours(geom, x, y, z)
    dead = np.zeros(len(x), np.uint8)
    merged = L.orc_cloud_merge(d, len(x), _ptr(x, _dp), _ptr(y, _dp), _ptr(z, _dp),
                               _ptr(b, _up), _ptr(neigh, _ip), _ptr(nn, _ip),
                               C.c_double(float(m["minDist"])), g.shape[1], _ptr(g, _dp),
                               _ptr(dead, _up))
    take(dead == 0)
    cap = 4096
    out = np.zeros(8 * cap)
    gout = np.zeros((cap, g.shape[1]))
    dom6 = np.array([m["xleft"], m["xright"], m["ybottom"], m["ytop"], m["zfront"],
                     m["zback"]])
    vlo, Nv = float(m["vmin"]), int(m["Nv"])
    dv = (float(m["vmax"]) - vlo) / (Nv - 1)
    added = L.orc_cloud_refill(d, len(x), _ptr(x, _dp), _ptr(y, _dp), _ptr(z, _dp),
                               _ptr(b, _up), _ptr(bid, _ip), _ptr(mom, _dp), nb, barr,
                               C.c_double(dx), _ptr(dom6, _dp), C.c_double(float(m["radius"])),
                               Nv, C.c_double(vlo), C.c_double(dv), C.c_double(float(m["R"])),
                               g.shape[1], cap, _ptr(out, _dp), _ptr(gout, _dp))
    assert added >= 0
    o = out.reshape(8, cap)[:, :added]
    x, y, z = [np.concatenate([a, o[k]]) for k, a in enumerate((x, y, z))]
    b = np.concatenate([b, np.zeros(added, np.uint8)])
    bid = np.concatenate([bid, np.full(added, -1, np.int32)])
    src = np.concatenate([src, np.full(added, -1, np.int32)])
    g = np.concatenate([g, gout[:added]])
    mom = np.concatenate([mom, o[3:8].T])
    return dict(x=x, y=y, z=z, boundary=b, bodyid=bid, g=g, mom=mom, src=src,
                counts=(removed, merged, added))
