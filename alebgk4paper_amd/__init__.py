"""alebgk4paper_amd -- the ALE-BGK time step on MI355X (gfx950).

Python mirror of the C ABI in include/alebgk_dev.h (libalebgk_hip.so) and of
the host case builder in include/alebgk_host.h (libalebgk_host.so). The
compute path is the HIP library; there is no CPU fallback: importing works
anywhere, but creating a Solver fails loudly when the library is missing or
no GPU is present.

    cs = configure("cavity3d", nx=40, nv=10)
    cloud = build_lattice(cs, precision=PREC_MIXED, mls_order=2)
    sol = Solver(cloud)
    sol.step(cs.dt)
    rho, ux, uy, uz, T = sol.moments()

Reference boundary replaced: Solver/Solver.cpp:67-89 (cudaUploadCloud,
cudaStep, cudaDownload*), Solver/CudaBackend.cu.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))
HIP_LIB = os.path.join(PKG, "libalebgk_hip.so")
HOST_LIB = os.path.join(PKG, "libalebgk_host.so")

MAX_NEIGH = 128
STATIC_SURF = 8192         # AB_STATIC_SURF: ab_body_loads_static's largest call
PREC_FP64, PREC_MIXED, PREC_FP64_TABLE = 0, 1, 2
PREC_NAMES = {PREC_FP64: "fp64", PREC_MIXED: "mixed", PREC_FP64_TABLE: "fp64-table"}
ERRORS = {1: "invalid argument", 2: "HIP runtime error", 3: "call out of order",
          4: "unsupported configuration", 5: "device allocation failed"}

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)
_up = C.POINTER(C.c_ubyte)
_fp = C.POINTER(C.c_float)


BODY_SQUARE, BODY_CIRCLE, BODY_SPHERE, BODY_CUBE = 1, 2, 3, 4


class ABBody(C.Structure):
    _fields_ = [("shape", C.c_int), ("c", C.c_double * 3), ("size", C.c_double),
                ("o", C.c_double * 2), ("rm", C.c_double * 9)]


def body(shape, centre, size, o=(1.0, 0.0), rm=(1, 0, 0, 0, 1, 0, 0, 0, 1)):
    """A rigid body for cloud management (RigidBody/Geometry.hpp): shape
    BODY_SQUARE (side size, orientation o), BODY_CIRCLE (radius), BODY_SPHERE
    (radius), BODY_CUBE (side size, body-to-world rotation rm, row-major)."""
    b = ABBody()
    b.shape = int(shape)
    b.c[:] = [float(v) for v in centre] + [0.0] * (3 - len(centre))
    b.size = float(size)
    b.o[:] = [float(v) for v in o]
    b.rm[:] = [float(v) for v in rm]
    return b


class ABRigidBody(C.Structure):
    """ab_rigid_body: a rigid body and its dynamic state (RigidBody,
    RigidBody/Geometry.hpp:19-80)."""
    _fields_ = [("shape", C.c_int), ("npts", C.c_int), ("rotates", C.c_int),
                ("size", C.c_double), ("mass", C.c_double), ("T", C.c_double),
                ("c", C.c_double * 3), ("o", C.c_double * 2), ("rm", C.c_double * 9),
                ("vel", C.c_double * 3), ("angular_vel", C.c_double), ("om", C.c_double * 3),
                ("force", C.c_double * 3), ("tq", C.c_double * 3)]

    def state(self):
        """The dynamic state as a dict of numpy arrays."""
        return dict(c=np.array(self.c[:]), vel=np.array(self.vel[:]),
                    om=np.array(self.om[:]), angular_vel=float(self.angular_vel),
                    o=np.array(self.o[:]), rm=np.array(self.rm[:]),
                    force=np.array(self.force[:]), tq=np.array(self.tq[:]))


class ABConfig(C.Structure):
    _fields_ = [("dim", C.c_int), ("n_particles", C.c_int), ("nv", C.c_int),
                ("vmin", C.c_double), ("vmax", C.c_double), ("R", C.c_double),
                ("tau", C.c_double), ("radius", C.c_double), ("mls_order", C.c_int),
                ("mls_tol", C.c_double), ("mls_clamp", C.c_int), ("no_mass_fix", C.c_int),
                ("precision", C.c_int), ("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int),
                ("xleft", C.c_double), ("ybottom", C.c_double), ("zfront", C.c_double),
                ("hbox", C.c_double), ("vbox", C.c_double), ("wbox", C.c_double),
                ("nbhbox", C.c_int), ("nbvbox", C.c_int), ("nbwbox", C.c_int),
                ("collision", C.c_int), ("dm_tol", C.c_double), ("dm_maxit", C.c_int),
                ("limiter", C.c_int)]


class ABCloudView(C.Structure):
    _fields_ = [("x", _dp), ("y", _dp), ("z", _dp), ("boundary", _up),
                ("wnx", _dp), ("wny", _dp), ("wnz", _dp), ("wux", _dp), ("wuy", _dp),
                ("wuz", _dp), ("wt", _dp), ("mvx", _dp), ("mvy", _dp), ("mvz", _dp),
                ("rho", _dp), ("ux", _dp), ("uy", _dp), ("uz", _dp), ("t", _dp),
                ("g", _dp), ("neigh", _ip), ("nneigh", _ip), ("body_id", _ip)]


class ABStepStats(C.Structure):
    _fields_ = [("mass_old", C.c_double), ("mass_after_transport", C.c_double),
                ("mass_after_collision", C.c_double), ("n_positivity", C.c_int),
                ("n_nad", C.c_int), ("n_mls_degenerate", C.c_int),
                ("n_mls_extrapolate", C.c_int), ("mls_ovr_sum", C.c_double),
                ("mls_ovr_max", C.c_double), ("dm_particles", C.c_int),
                ("dm_unconverged", C.c_int), ("dm_passes", C.c_longlong),
                ("dm_res_max", C.c_double), ("n_pad_fail", C.c_int)]


class ALHCase(C.Structure):
    _fields_ = [("id", C.c_int), ("name", C.c_char * 32), ("dim", C.c_int),
                ("has_body", C.c_int), ("lid", C.c_int), ("lid_speed", C.c_double),
                ("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int), ("nv", C.c_int),
                ("vmin", C.c_double), ("vmax", C.c_double), ("dt", C.c_double),
                ("tfinal", C.c_double), ("tau", C.c_double), ("rho0", C.c_double),
                ("Tw", C.c_double), ("R", C.c_double), ("d_hs", C.c_double),
                ("save_every", C.c_int), ("mls_order", C.c_int), ("dom", C.c_double * 6)]


class CkptInfo(C.Structure):
    _fields_ = [("dim", C.c_int), ("n", C.c_longlong), ("gstride", C.c_longlong),
                ("step", C.c_longlong), ("t", C.c_double), ("nbodies", C.c_int)]


class CkptFields(C.Structure):
    _fields_ = [("x", _dp), ("y", _dp), ("z", _dp), ("mom", _dp), ("mesh", _dp), ("wall", _dp),
                ("grid", _dp), ("ints", _ip), ("flags", _up), ("g", _dp), ("bodies", _dp)]


CKPT_SHAPES = {"x": None, "y": None, "z": None, "mom": 6, "mesh": 3, "wall": 7, "grid": 3,
               "ints": 4, "flags": 3}


class ALEBGKError(RuntimeError):
    pass


def _ptr(a, t):
    if a is None:
        return C.cast(None, t)
    if not a.flags["C_CONTIGUOUS"]:
        raise ValueError("arrays passed to the library must be C-contiguous")
    return a.ctypes.data_as(t)


_hip = None
_host = None

HIP_EXPORTS = ["ab_create", "ab_destroy", "ab_last_error", "ab_upload", "ab_rebuild_neighbours",
               "ab_download_neighbours", "ab_step", "ab_steps", "ab_transport_only",
               "ab_download_g", "ab_download_gt", "ab_download_moments", "ab_total_mass",
               "ab_body_loads", "ab_set_wall_state", "ab_synchronize", "ab_profile",
               "ab_profile_read", "ab_stage_name", "ab_dm_download", "ab_dm_record",
               "ab_dm_rows", "ab_mlp_init", "ab_mlp_train", "ab_mlp_enable", "ab_mlp_get",
               "ab_mlp_set", "ab_mlp_forward", "ab_mlp_infer", "ab_mlp_loss_grad", "ab_graph_create",
               "ab_graph_launch", "ab_graph_synchronize", "ab_graph_destroy", "ab_autotune",
               "ab_set_transport_shape", "ab_copy_calibration", "ab_download_voxels",
               "ab_cloud_manage", "ab_cloud_source_map", "ab_download_particles",
               "ab_batch_create", "ab_batch_scatter", "ab_check_config", "ab_download_body_ids",
               "ab_body_step", "ab_body_loads_static", "ab_body_step_many"]
MLP_TRAINABLE = 4869
MLP_PARAMS = 4889
COLLISION_CONTINUOUS = 0
COLLISION_DISCRETE = 1
LIMITER_GPU = 0          # the reference's GPU step (cudaStep)
LIMITER_MOOD = 1         # its CPU step with the MOOD limiter (runSolver)
HOST_EXPORTS = ["alh_configure", "alh_lattice_size", "alh_build_lattice", "alh_ckpt_info_read",
                "alh_ckpt_read", "alh_ckpt_write", "alh_make_bodies", "alh_body_case_size",
                "alh_build_body_case", "alh_diagnostics", "alh_write_vtk", "alh_write_timeseries",
                "alh_format_step", "alh_write_bodies", "alh_guard_init", "alh_guard_check"]


def hip_lib():
    """libalebgk_hip.so; raises if it has not been built."""
    global _hip
    if _hip is None:
        if not os.path.exists(HIP_LIB):
            raise ALEBGKError("libalebgk_hip.so is not built: run __graft_entry__.build() "
                              "or python -m alebgk4paper_amd.build")
        L = C.CDLL(HIP_LIB)
        L.ab_last_error.restype = C.c_char_p
        L.ab_stage_name.restype = C.c_char_p
        L.ab_step.argtypes = [C.c_void_p, C.c_double, C.c_void_p]
        L.ab_steps.argtypes = [C.c_void_p, C.c_double, C.c_int]
        L.ab_transport_only.argtypes = [C.c_void_p, C.c_double]
        L.ab_mlp_train.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_float, C.c_ulonglong,
                                   C.c_void_p, C.c_void_p]
        _hip = L
    return _hip


def host_lib():
    global _host
    if _host is None:
        if not os.path.exists(HOST_LIB):
            raise ALEBGKError("libalebgk_host.so is not built")
        _host = C.CDLL(HOST_LIB)
    return _host


@dataclass
class Case:
    """One reference case after overrides (TestCases.cpp:113-228)."""
    raw: ALHCase

    def __getattr__(self, k):
        v = getattr(self.raw, k)
        return v.decode() if isinstance(v, bytes) else v

    @property
    def dom(self):
        return tuple(self.raw.dom)


def configure(token, nx=0, nv=0, mls_order=None):
    cs = ALHCase()
    rc = host_lib().alh_configure(str(token).encode(), C.c_int(nx), C.c_int(nv), C.byref(cs))
    if rc:
        raise ALEBGKError(f"unknown case {token!r}")
    if mls_order is not None:
        cs.mls_order = int(mls_order)
    return Case(cs)


@dataclass
class Cloud:
    """Host arrays of one cloud plus its ab_config."""
    cfg: ABConfig
    x: np.ndarray
    y: np.ndarray
    z: np.ndarray
    boundary: np.ndarray
    wall: np.ndarray          # N x 7: wn(3), wU(3), wT
    mom: np.ndarray           # N x 5: rho, ux, uy, uz, T
    g: np.ndarray             # N x gstride
    neigh: np.ndarray | None = None
    nneigh: np.ndarray | None = None
    mesh: np.ndarray | None = None   # N x 3 ALE mesh velocity
    record: dict | None = None       # restart-file fields not held elsewhere (p, bodyId, flags)
    body_id: np.ndarray | None = None   # Particle::bodyId (-1: none)
    bodies: list | None = None       # ABRigidBody of a body case
    geo: tuple | None = None         # dx, dy, dz, minDist of the lattice

    @property
    def N(self):
        return len(self.x)

    @property
    def gstride(self):
        return self.g.shape[1]


def build_lattice(case, precision=PREC_FP64, mls_order=None):
    """generateParticles + initWallData + applyLidProfile + applyIC."""
    cs = case.raw
    if mls_order is not None:
        cs.mls_order = int(mls_order)
    n = C.c_int()
    host_lib().alh_lattice_size(C.byref(cs), C.byref(n))
    N = n.value
    gs = cs.nv ** 3 if cs.dim == 3 else 2 * cs.nv ** 2
    x, y, z = np.zeros(N), np.zeros(N), np.zeros(N)
    b = np.zeros(N, np.uint8)
    wall = np.zeros((N, 7))
    mom = np.zeros((N, 5))
    g = np.zeros((N, gs))
    cfg = ABConfig()
    rc = host_lib().alh_build_lattice(C.byref(cs), C.c_int(precision), C.byref(cfg),
                                      _ptr(x, _dp), _ptr(y, _dp), _ptr(z, _dp), _ptr(b, _up),
                                      _ptr(wall, _dp), _ptr(mom, _dp), _ptr(g, _dp))
    if rc:
        raise ALEBGKError(f"alh_build_lattice: {ERRORS.get(rc, rc)}")
    return Cloud(cfg, x, y, z, b, wall, mom, g)


def make_bodies(case, body_scale=1.0):
    """makeBodies (TestCases.cpp:251-294): the case's bodies at rest."""
    arr = (ABRigidBody * 4)()
    n = C.c_int()
    rc = host_lib().alh_make_bodies(C.byref(case.raw), C.c_double(body_scale), arr, C.byref(n))
    if rc:
        raise ALEBGKError(f"alh_make_bodies: {ERRORS.get(rc, rc)}")
    return [arr[k] for k in range(n.value)]


def build_case(case, precision=PREC_FP64, mls_order=None, body_scale=1.0):
    """The cloud of any case: build_lattice for the plain cavities, the
    embedded body case (generateParticles + embedRigidBodies + wall state +
    parabolic lid + applyIC) for the others. A body cloud carries body_id,
    bodies and geo (dx, dy, dz, minDist)."""
    cs = case.raw
    if mls_order is not None:
        cs.mls_order = int(mls_order)
    if not cs.has_body:
        return build_lattice(case, precision=precision, mls_order=mls_order)
    n = C.c_int()
    rc = host_lib().alh_body_case_size(C.byref(cs), C.c_double(body_scale), C.byref(n))
    if rc:
        raise ALEBGKError(f"alh_body_case_size: {ERRORS.get(rc, rc)}")
    N = n.value
    gs = cs.nv ** 3 if cs.dim == 3 else 2 * cs.nv ** 2
    x, y, z = np.zeros(N), np.zeros(N), np.zeros(N)
    b = np.zeros(N, np.uint8)
    wall = np.zeros((N, 7))
    mom = np.zeros((N, 5))
    g = np.zeros((N, gs))
    bid = np.zeros(N, np.int32)
    arr = (ABRigidBody * 4)()
    nb = C.c_int()
    geo = np.zeros(4)
    cfg = ABConfig()
    rc = host_lib().alh_build_body_case(C.byref(cs), C.c_double(body_scale), C.c_int(precision),
                                        C.byref(cfg), _ptr(x, _dp), _ptr(y, _dp), _ptr(z, _dp),
                                        _ptr(b, _up), _ptr(wall, _dp), _ptr(mom, _dp),
                                        _ptr(g, _dp), _ptr(bid, _ip), arr, C.byref(nb),
                                        _ptr(geo, _dp))
    if rc:
        raise ALEBGKError(f"alh_build_body_case: {ERRORS.get(rc, rc)}")
    return Cloud(cfg, x, y, z, b, wall, mom, g, body_id=bid,
                 bodies=[arr[k] for k in range(nb.value)], geo=tuple(geo))


class Diag(C.Structure):
    """alh_diag: computeDiagnostics' means over the interior particles."""
    _fields_ = [("rho_avg", C.c_double), ("KE", C.c_double), ("T_avg", C.c_double),
                ("p_avg", C.c_double), ("mood_count", C.c_int), ("n", C.c_int)]


def diagnostics(boundary, rho, ux, uy, uz, T, p, mood):
    """computeDiagnostics (Core/Output.hpp:65-84) through alh_diagnostics."""
    a = [np.ascontiguousarray(v, dtype=np.float64) for v in (rho, ux, uy, uz, T, p)]
    b = np.ascontiguousarray(boundary, dtype=np.uint8)
    m = np.ascontiguousarray(mood, dtype=np.uint8)
    d = Diag()
    rc = host_lib().alh_diagnostics(C.c_int(len(b)), _ptr(b, _up), *[_ptr(v, _dp) for v in a],
                                    _ptr(m, _up), C.byref(d))
    if rc:
        raise ALEBGKError(f"alh_diagnostics: {ERRORS.get(rc, rc)}")
    return d


def write_vtk(out_dir, step, t, x, y, z, rho, T, p, nv, mood, boundary, ux, uy, uz):
    """writeVTK (Core/Output.hpp:17-51): out_dir/output_<step>.vtk."""
    f8 = [np.ascontiguousarray(v, dtype=np.float64) for v in (x, y, z, rho, T, p)]
    u8 = [np.ascontiguousarray(v, dtype=np.uint8) for v in (mood, boundary)]
    uu = [np.ascontiguousarray(v, dtype=np.float64) for v in (ux, uy, uz)]
    rc = host_lib().alh_write_vtk(str(out_dir).encode(), C.c_int(step), C.c_double(t),
                                  C.c_int(len(f8[0])), *[_ptr(v, _dp) for v in f8], C.c_int(nv),
                                  *[_ptr(v, _up) for v in u8], *[_ptr(v, _dp) for v in uu])
    if rc:
        raise ALEBGKError(f"alh_write_vtk: {ERRORS.get(rc, rc)}")


def write_timeseries(path, first, t, d):
    """writeTimeseries (Core/Output.hpp:89-106): one CSV row."""
    rc = host_lib().alh_write_timeseries(str(path).encode(), C.c_int(int(first)), C.c_double(t),
                                         C.byref(d))
    if rc:
        raise ALEBGKError(f"alh_write_timeseries: {ERRORS.get(rc, rc)}")


def format_step(step, t, dt, d):
    """printStep's line (Core/Output.hpp:111-115)."""
    buf = C.create_string_buffer(160)
    host_lib().alh_format_step(C.c_int(step), C.c_double(t), C.c_double(dt), C.byref(d), buf,
                               C.c_int(len(buf)))
    return buf.value.decode()


def write_bodies(path, first, t, dim, bodies):
    """writeBodyLog (Solver/Solver.cpp:96-121): one bodies.csv row per body."""
    arr = (ABRigidBody * max(len(bodies), 1))(*bodies)
    rc = host_lib().alh_write_bodies(str(path).encode(), C.c_int(int(first)), C.c_double(t),
                                     C.c_int(dim), C.c_int(len(bodies)), arr)
    if rc:
        raise ALEBGKError(f"alh_write_bodies: {ERRORS.get(rc, rc)}")


class Guard(C.Structure):
    """stability::Guard (Core/Stability.hpp:46-123) through alh_guard_*: the
    bounds and the ALEBGK_GUARD_* / ALEBGK_NO_GUARD overrides are read when
    it is made."""
    _fields_ = [("Tw", C.c_double), ("rho0", C.c_double), ("ke_ref", C.c_double),
                ("have_ke", C.c_int), ("enabled", C.c_int), ("t_lo", C.c_double),
                ("t_hi", C.c_double), ("rho_lo", C.c_double), ("rho_hi", C.c_double),
                ("ke_hi", C.c_double)]

    def __init__(self, wall_temp, rho0):
        super().__init__()
        host_lib().alh_guard_init(C.byref(self), C.c_double(wall_temp), C.c_double(rho0))

    def diverged(self, step, t, d):
        """(True, reason) once the diagnostics leave the bounds."""
        buf = C.create_string_buffer(256)
        r = host_lib().alh_guard_check(C.byref(self), C.c_int(step), C.c_double(t), C.byref(d),
                                       buf, C.c_int(len(buf)))
        return bool(r), buf.value.decode()


class ABBodyTask(C.Structure):
    _fields_ = [("dt", C.c_double), ("nbodies", C.c_int), ("bodies", C.POINTER(ABRigidBody)),
                ("dx", C.c_double), ("min_dist", C.c_double), ("dom", C.c_double * 6),
                ("loads", _dp), ("n_new", _ip), ("counts", _ip)]


def body_step_many(solvers, dt, bodies, geos, doms):
    """ab_body_step_many: the moving-body block of several solvers at once
    (one host thread per cloud). bodies[k]: solver k's ABRigidBody list,
    updated in place; geos[k]: its (dx, dy, dz, minDist); doms[k]: its box as
    box() gives it. Returns the (removed, merged, added) counts per solver."""
    n = len(solvers)
    arrs = [(ABRigidBody * max(len(b), 1))(*b) for b in bodies]
    n_new = (C.c_int * n)()
    counts = (C.c_int * (3 * n))()
    tasks = (ABBodyTask * n)()
    for k in range(n):
        t = tasks[k]
        t.dt = float(dt)
        t.nbodies = len(bodies[k])
        t.bodies = C.cast(arrs[k], C.POINTER(ABRigidBody))
        t.dx, t.min_dist = float(geos[k][0]), float(geos[k][3])
        t.dom[:] = [float(v) for v in doms[k]]
        t.loads = C.cast(None, _dp)
        t.n_new = C.cast(C.byref(n_new, 4 * k), _ip)
        t.counts = C.cast(C.byref(counts, 12 * k), _ip)
    ctxs = (C.c_void_p * n)(*[s._ctx.value for s in solvers])
    rc = hip_lib().ab_body_step_many(C.c_int(n), ctxs, tasks)
    if rc:
        msgs = "; ".join(s._lib.ab_last_error(s._ctx).decode() for s in solvers)
        raise ALEBGKError(f"ab_body_step_many: {ERRORS.get(rc, rc)}: {msgs}")
    for k in range(n):
        for j in range(len(bodies[k])):
            C.memmove(C.byref(bodies[k][j]), C.byref(arrs[k][j]), C.sizeof(ABRigidBody))
        solvers[k].N = n_new[k]
    return [tuple(counts[3 * k:3 * k + 3]) for k in range(n)]


def box(case):
    """The case's box as ab_cloud_manage and ab_body_step take it: (xleft,
    xright, ybottom, ytop, zfront, zback)."""
    d = case.dom
    return (d[0], d[1], d[3], d[2], d[4], d[5])


class Solver:
    """One device context (ab_ctx): the cloud resident on the GPU."""

    def __init__(self, cloud: Cloud, precision=None, mls_order=None, no_mass_fix=False,
                 mls_clamp=False, collision=None, dm_tol=None, dm_maxit=None, limiter=None):
        self._lib = hip_lib()
        cfg = ABConfig()
        C.memmove(C.byref(cfg), C.byref(cloud.cfg), C.sizeof(ABConfig))
        cfg.n_particles = cloud.N
        if precision is not None:
            cfg.precision = int(precision)
        if mls_order is not None:
            cfg.mls_order = int(mls_order)
        cfg.no_mass_fix = int(bool(no_mass_fix))
        cfg.mls_clamp = int(bool(mls_clamp))
        if collision is not None:
            cfg.collision = int(collision)
        if dm_tol is not None:
            cfg.dm_tol = float(dm_tol)
        if dm_maxit is not None:
            cfg.dm_maxit = int(dm_maxit)
        if limiter is not None:
            cfg.limiter = int(limiter)
        self.cfg = cfg
        self.cloud = cloud
        self.N, self.dim = cloud.N, cfg.dim
        self.gstride = cloud.gstride
        self._ctx = C.c_void_p()
        rc = self._lib.ab_create(C.byref(cfg), C.byref(self._ctx))
        if rc:
            raise ALEBGKError(f"ab_create: {ERRORS.get(rc, rc)}")
        self.upload(cloud)

    def _check(self, rc, what):
        if rc:
            msg = self._lib.ab_last_error(self._ctx)
            raise ALEBGKError(f"{what}: {ERRORS.get(rc, rc)}: {msg.decode() if msg else ''}")

    def upload(self, cloud: Cloud):
        f8 = lambda a: np.ascontiguousarray(a, dtype=np.float64)
        keep = dict(x=f8(cloud.x), y=f8(cloud.y), z=f8(cloud.z),
                    b=np.ascontiguousarray(cloud.boundary, dtype=np.uint8),
                    w=[f8(cloud.wall[:, k]) for k in range(7)],
                    m=[f8(cloud.mom[:, k]) for k in range(5)], g=f8(cloud.g))
        v = ABCloudView()
        v.x, v.y, v.z = _ptr(keep["x"], _dp), _ptr(keep["y"], _dp), _ptr(keep["z"], _dp)
        v.boundary = _ptr(keep["b"], _up)
        v.wnx, v.wny, v.wnz, v.wux, v.wuy, v.wuz, v.wt = [_ptr(a, _dp) for a in keep["w"]]
        if cloud.mesh is not None:
            keep["mv"] = [f8(cloud.mesh[:, k]) for k in range(3)]
            v.mvx, v.mvy, v.mvz = [_ptr(a, _dp) for a in keep["mv"]]
        else:
            v.mvx = v.mvy = v.mvz = C.cast(None, _dp)
        v.rho, v.ux, v.uy, v.uz, v.t = [_ptr(a, _dp) for a in keep["m"]]
        v.g = _ptr(keep["g"], _dp)
        if cloud.body_id is not None:
            keep["bid"] = np.ascontiguousarray(cloud.body_id, dtype=np.int32)
            v.body_id = _ptr(keep["bid"], _ip)
        else:
            v.body_id = C.cast(None, _ip)
        if cloud.neigh is not None:
            keep["nb"] = np.ascontiguousarray(cloud.neigh, dtype=np.int32)
            keep["nn"] = np.ascontiguousarray(cloud.nneigh, dtype=np.int32)
            v.neigh, v.nneigh = _ptr(keep["nb"], _ip), _ptr(keep["nn"], _ip)
        else:
            v.neigh, v.nneigh = C.cast(None, _ip), C.cast(None, _ip)
        self._check(self._lib.ab_upload(self._ctx, C.byref(v)), "ab_upload")

    def close(self):
        if self._ctx:
            self._lib.ab_destroy(self._ctx)
            self._ctx = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- stepping --------------------------------------------------------
    def step(self, dt, stats=False):
        if stats:
            st = ABStepStats()
            self._check(self._lib.ab_step(self._ctx, dt, C.byref(st)), "ab_step")
            return {k: getattr(st, k) for k, _ in ABStepStats._fields_}
        self._check(self._lib.ab_step(self._ctx, dt, None), "ab_step")
        return None

    def steps(self, dt, n):
        self._check(self._lib.ab_steps(self._ctx, dt, int(n)), "ab_steps")

    def transport_only(self, dt):
        self._check(self._lib.ab_transport_only(self._ctx, dt), "ab_transport_only")

    def synchronize(self):
        self._check(self._lib.ab_synchronize(self._ctx), "ab_synchronize")

    def rebuild_neighbours(self):
        self._check(self._lib.ab_rebuild_neighbours(self._ctx), "ab_rebuild_neighbours")

    # ---- restart files ----------------------------------------------------
    def voxels(self):
        v = np.zeros(self.N, np.int32)
        self._check(self._lib.ab_download_voxels(self._ctx, _ptr(v, _ip)), "ab_download_voxels")
        return v

    def save_checkpoint(self, path, step, t):
        """ckpt::save of the device state: g and rho, u, T from the GPU,
        p = rho R T, the geometry and wall state of the uploaded cloud, the
        voxels of the device rebuild; fields a restart file carried over
        (bodyId, validg, moodFlag, bodies) are kept."""
        cl, cfg = self.cloud, self.cfg
        N = self.N
        mom = self.moments()
        p = mom[:, 0] * cfg.R * mom[:, 4]
        rec = cl.record or {}
        dv = (cfg.vmax - cfg.vmin) / (cfg.nv - 1)
        grid = rec.get("grid", np.tile([cfg.vmin, cfg.vmax, dv], (N, 1)))
        ints = np.array(rec["ints"]) if "ints" in rec else \
            np.column_stack([np.full(N, cfg.nv), np.full(N, cfg.dim), np.full(N, -1),
                             np.zeros(N)]).astype(np.int32)
        ints[:, 3] = self.voxels()
        flags = np.array(rec["flags"]) if "flags" in rec else \
            np.column_stack([cl.boundary, np.ones(N), np.zeros(N)]).astype(np.uint8)
        flags[:, 0] = cl.boundary
        fields = {"x": cl.x, "y": cl.y, "z": cl.z, "mom": np.column_stack([mom, p]),
                  "mesh": cl.mesh if cl.mesh is not None else np.zeros((N, 3)),
                  "wall": cl.wall, "grid": grid, "ints": ints, "flags": flags, "g": self.g(),
                  "bodies": rec.get("bodies", np.zeros((0, 30)))}
        checkpoint_write(path, step, t, fields, dim=cfg.dim)

    # ---- transport launch shape ------------------------------------------
    def autotune(self, dt, reps=3):
        """Times every transport launch shape that fits the lattice, keeps the
        fastest; returns ([(index, name, ms)], chosen index)."""
        cap = 32
        names = ((C.c_char * 32) * cap)()
        ms = (C.c_float * cap)()
        n, ch = C.c_int(0), C.c_int(-1)
        self._check(self._lib.ab_autotune(self._ctx, C.c_double(dt), int(reps), cap, names, ms,
                                          C.byref(n), C.byref(ch)), "ab_autotune")
        table = [(i, names[i].value.decode(), ms[i]) for i in range(n.value) if ms[i] >= 0]
        return table, ch.value

    def set_transport_shape(self, index):
        self._check(self._lib.ab_set_transport_shape(self._ctx, int(index)), "ab_set_transport_shape")

    # ---- discrete-Maxwellian collision and the MLP warm start --------------
    def dm_download(self):
        """(alpha N x 5, passes N, residual N) of the last step."""
        a = np.zeros((self.N, 5)); p = np.zeros(self.N, np.int32); r = np.zeros(self.N)
        self._check(self._lib.ab_dm_download(self._ctx, _ptr(a, _dp), _ptr(p, _ip), _ptr(r, _dp)),
                    "ab_dm_download")
        return a, p, r

    def dm_record(self, steps):
        self._check(self._lib.ab_dm_record(self._ctx, int(steps)), "ab_dm_record")

    def dm_rows(self):
        n, v = C.c_int(0), C.c_int(0)
        self._check(self._lib.ab_dm_rows(self._ctx, C.byref(n), C.byref(v)), "ab_dm_rows")
        return n.value, v.value

    def mlp_init(self, seed=1):
        self._check(self._lib.ab_mlp_init(self._ctx, C.c_ulonglong(seed)), "ab_mlp_init")

    def mlp_train(self, iters, batch=4096, lr=1e-3, seed=7):
        a, b = C.c_float(0), C.c_float(0)
        self._check(self._lib.ab_mlp_train(self._ctx, int(iters), int(batch), C.c_float(lr),
                                           C.c_ulonglong(seed), C.byref(a), C.byref(b)),
                    "ab_mlp_train")
        return a.value, b.value

    def mlp_enable(self, on=True):
        self._check(self._lib.ab_mlp_enable(self._ctx, int(bool(on))), "ab_mlp_enable")

    def mlp_get(self):
        p = np.zeros(MLP_PARAMS, np.float32)
        self._check(self._lib.ab_mlp_get(self._ctx, _ptr(p, _fp)), "ab_mlp_get")
        return p

    def mlp_set(self, params):
        p = np.ascontiguousarray(params, dtype=np.float32)
        assert p.size == MLP_PARAMS
        self._check(self._lib.ab_mlp_set(self._ctx, _ptr(p, _fp)), "ab_mlp_set")

    def mlp_forward(self, x):
        x = np.ascontiguousarray(x, dtype=np.float32)
        y = np.zeros((x.shape[0], 5), np.float32)
        self._check(self._lib.ab_mlp_forward(self._ctx, _ptr(x, _fp), x.shape[0], _ptr(y, _fp)),
                    "ab_mlp_forward")
        return y

    def mlp_infer(self, x):
        """The warm start's inference as the step runs it (BF16 weights on
        MFMA, BF16 layer-1 activations, FP32 accumulation)."""
        x = np.ascontiguousarray(x, dtype=np.float32)
        y = np.zeros((x.shape[0], 5), np.float32)
        self._check(self._lib.ab_mlp_infer(self._ctx, _ptr(x, _fp), x.shape[0], _ptr(y, _fp)),
                    "ab_mlp_infer")
        return y

    def mlp_loss_grad(self, x, t, w):
        x = np.ascontiguousarray(x, dtype=np.float32)
        t = np.ascontiguousarray(t, dtype=np.float32)
        w = np.ascontiguousarray(w, dtype=np.float32)
        loss = C.c_float(0)
        grad = np.zeros(MLP_TRAINABLE, np.float32)
        self._check(self._lib.ab_mlp_loss_grad(self._ctx, _ptr(x, _fp), _ptr(t, _fp), _ptr(w, _fp),
                                               x.shape[0], C.byref(loss), _ptr(grad, _fp)),
                    "ab_mlp_loss_grad")
        return loss.value, grad

    # ---- downloads --------------------------------------------------------
    def g(self):
        out = np.zeros((self.N, self.gstride))
        self._check(self._lib.ab_download_g(self._ctx, _ptr(out, _dp)), "ab_download_g")
        return out

    def gt(self):
        out = np.zeros((self.N, self.gstride))
        self._check(self._lib.ab_download_gt(self._ctx, _ptr(out, _dp)), "ab_download_gt")
        return out

    def moments(self, with_limflag=False):
        a = [np.zeros(self.N) for _ in range(6)]
        self._check(self._lib.ab_download_moments(self._ctx, *[_ptr(x, _dp) for x in a]),
                    "ab_download_moments")
        mom = np.stack(a[:5], axis=1)
        return (mom, a[5]) if with_limflag else mom

    def neighbours(self):
        nb = np.zeros((self.N, MAX_NEIGH), np.int32)
        nn = np.zeros(self.N, np.int32)
        self._check(self._lib.ab_download_neighbours(self._ctx, _ptr(nb, _ip), _ptr(nn, _ip)),
                    "ab_download_neighbours")
        return nb, nn

    def total_mass(self):
        m = C.c_double()
        self._check(self._lib.ab_total_mass(self._ctx, C.byref(m)), "ab_total_mass")
        return m.value

    def particles(self):
        """Positions, boundary flags and wall state of the device cloud
        (ab_download_particles): dict x, y, z, boundary, wall (N x 7)."""
        x, y, z = np.zeros(self.N), np.zeros(self.N), np.zeros(self.N)
        b = np.zeros(self.N, np.uint8)
        w = np.zeros((self.N, 7))
        self._check(self._lib.ab_download_particles(self._ctx, _ptr(x, _dp), _ptr(y, _dp),
                                                    _ptr(z, _dp), _ptr(b, _up), _ptr(w, _dp)),
                    "ab_download_particles")
        return dict(x=x, y=y, z=z, boundary=b, wall=w)

    def cloud_manage(self, bodies, body_id, dx, min_dist, dom):
        """manageRemove, mergeCloseParticles and manageRefill with the
        neighbour rebuilds between them (RigidBody/CloudManagement.cpp:17-275,
        Solver/Solver.cpp:379-398), on the device (ab_cloud_manage). bodies:
        ABBody values (see body()); body_id: Particle::bodyId of the current
        cloud; dom: (xleft, xright, ybottom, ytop, zfront, zback). Returns
        (src, counts): for each particle of the new cloud its index before
        the call (-1 for a refilled one), and (removed, merged, added)."""
        nb = len(bodies)
        arr = (ABBody * max(nb, 1))(*bodies)
        bid = np.ascontiguousarray(body_id, dtype=np.int32)
        if len(bid) != self.N:
            raise ValueError("body_id needs one entry per particle")
        d6 = np.ascontiguousarray(dom, dtype=np.float64)
        n_new = C.c_int()
        counts = (C.c_int * 3)()
        self._check(self._lib.ab_cloud_manage(self._ctx, C.c_int(nb), arr, _ptr(bid, _ip),
                                              C.c_double(dx), C.c_double(min_dist),
                                              _ptr(d6, _dp), C.byref(n_new), counts),
                    "ab_cloud_manage")
        self.N = n_new.value
        src = np.zeros(self.N, np.int32)
        self._check(self._lib.ab_cloud_source_map(self._ctx, _ptr(src, _ip)),
                    "ab_cloud_source_map")
        return src, tuple(counts)

    def body_ids(self):
        """Particle::bodyId of the device cloud (ab_download_body_ids)."""
        b = np.zeros(self.N, np.int32)
        self._check(self._lib.ab_download_body_ids(self._ctx, _ptr(b, _ip)), "ab_download_body_ids")
        return b

    def body_step(self, dt, bodies, dx, min_dist, dom):
        """The moving-body block after a step (ab_body_step): Heun-coupled
        loads, Newton-Euler, contact, surface wall state, cloud management.
        bodies: ABRigidBody list, updated in place. dom: (xleft, xright,
        ybottom, ytop, zfront, zback). Returns (loads (nb x 12: L0 force,
        torque, L1 force, torque), counts (removed, merged, added))."""
        nb = len(bodies)
        arr = (ABRigidBody * max(nb, 1))(*bodies)
        d6 = np.ascontiguousarray(dom, dtype=np.float64)
        ld = np.zeros(12 * max(nb, 1))
        counts = (C.c_int * 3)()
        n_new = C.c_int()
        self._check(self._lib.ab_body_step(self._ctx, C.c_double(dt), C.c_int(nb), arr,
                                           C.c_double(dx), C.c_double(min_dist), _ptr(d6, _dp),
                                           _ptr(ld, _dp), C.byref(n_new), counts), "ab_body_step")
        for k in range(nb):
            C.memmove(C.byref(bodies[k]), C.byref(arr[k]), C.sizeof(ABRigidBody))
        self.N = n_new.value
        return ld.reshape(-1, 12)[:nb], tuple(counts)

    def set_wall_state(self, idx, xyz=None, wn=None, wu=None, mv=None):
        """Positions, wall normals, wall velocities and mesh velocities of
        the particles idx (the uploads of cudaSyncBodies,
        Solver/CudaBackend.cu:1884-1915; ab_set_wall_state). Each array is
        len(idx) x 3, or None to leave that field."""
        ii = np.ascontiguousarray(idx, dtype=np.int32)
        n = len(ii)

        def a3(a):
            if a is None:
                return None
            a = np.ascontiguousarray(a, dtype=np.float64).reshape(n, 3)
            return a
        xyz, wn, wu, mv = a3(xyz), a3(wn), a3(wu), a3(mv)
        cols = [None, None, None] if xyz is None else [np.ascontiguousarray(xyz[:, k])
                                                       for k in range(3)]
        self._check(self._lib.ab_set_wall_state(
            self._ctx, C.c_int(n), _ptr(ii, _ip), _ptr(cols[0], _dp), _ptr(cols[1], _dp),
            _ptr(cols[2], _dp), _ptr(wn, _dp), _ptr(wu, _dp), _ptr(mv, _dp)),
            "ab_set_wall_state")

    def body_loads(self, surf_idx, surf_body, centres, areas, wu=None):
        si = np.ascontiguousarray(surf_idx, dtype=np.int32)
        sb = np.ascontiguousarray(surf_body, dtype=np.int32)
        bc = np.ascontiguousarray(centres, dtype=np.float64).reshape(-1)
        ba = np.ascontiguousarray(areas, dtype=np.float64).reshape(-1)
        wuu = None if wu is None else np.ascontiguousarray(wu, dtype=np.float64).reshape(-1)
        nb = len(ba)
        out = np.zeros(6 * nb)
        self._check(self._lib.ab_body_loads(self._ctx, C.c_int(len(si)), _ptr(si, _ip),
                                            _ptr(sb, _ip), C.c_int(nb), _ptr(bc, _dp),
                                            _ptr(ba, _dp), _ptr(wuu, _dp), _ptr(out, _dp)),
                    "ab_body_loads")
        return out.reshape(nb, 6)

    # ---- profiling ------------------------------------------------------------
    def profile(self, enable=True):
        self._check(self._lib.ab_profile(self._ctx, C.c_int(int(enable))), "ab_profile")

    def profile_read(self):
        ms = (C.c_double * 16)()
        cnt = (C.c_long * 16)()
        ns = C.c_int()
        self._check(self._lib.ab_profile_read(self._ctx, ms, cnt, C.byref(ns)), "ab_profile_read")
        return {self._lib.ab_stage_name(k).decode(): (ms[k], cnt[k]) for k in range(ns.value)}


# ---- restart files (the reference's checkpoint format v1) -------------------

def cloud_from_checkpoint(path, case, precision=PREC_FP64, mls_order=None):
    """A cloud configured for `case` (lattice, voxel grid, neighbour radius)
    with every field the restart file holds restored. The neighbour lists are
    left to the device rebuild at upload, as the reference rebuilds them on
    load (Solver/Solver.cpp:281)."""
    base = build_lattice(case, precision, mls_order)
    d = checkpoint_read(path, n=base.N, gstride=base.gstride)
    rec = {"p": d["mom"][:, 5].copy(), "ints": d["ints"].copy(), "flags": d["flags"].copy(),
           "grid": d["grid"].copy(), "bodies": d["bodies"].copy(), "step": d["step"], "t": d["t"]}
    return Cloud(base.cfg, d["x"], d["y"], d["z"], d["flags"][:, 0].copy(), d["wall"],
                 d["mom"][:, :5].copy(), d["g"], mesh=d["mesh"], record=rec)

def checkpoint_info(path):
    info = CkptInfo()
    rc = host_lib().alh_ckpt_info_read(str(path).encode(), C.byref(info))
    if rc:
        raise ALEBGKError(f"alh_ckpt_info_read({path}): {ERRORS.get(rc, rc)}")
    return {k: getattr(info, k) for k, _ in CkptInfo._fields_}


def _ckpt_fields(d):
    f = CkptFields()
    for k in ("x", "y", "z", "mom", "mesh", "wall", "grid", "g", "bodies"):
        setattr(f, k, _ptr(d[k], _dp) if d.get(k) is not None and d[k].size else C.cast(None, _dp))
    f.ints = _ptr(d["ints"], _ip)
    f.flags = _ptr(d["flags"], _up)
    return f


def checkpoint_read(path, n=None, gstride=None, nbodies=None):
    """The record fields of a restart file as arrays (mom: rho ux uy uz T p;
    ints: Nv_local dim bodyId voxel; flags: boundary validg moodFlag). The
    file is refused if it does not match the given n / gstride / nbodies."""
    info = checkpoint_info(path)
    exp = CkptInfo(info["dim"], info["n"] if n is None else n,
                   info["gstride"] if gstride is None else gstride, info["step"], info["t"],
                   info["nbodies"] if nbodies is None else nbodies)
    N, gs, nb = exp.n, exp.gstride, exp.nbodies
    d = {"x": np.zeros(N), "y": np.zeros(N), "z": np.zeros(N), "mom": np.zeros((N, 6)),
         "mesh": np.zeros((N, 3)), "wall": np.zeros((N, 7)), "grid": np.zeros((N, 3)),
         "ints": np.zeros((N, 4), np.int32), "flags": np.zeros((N, 3), np.uint8),
         "g": np.zeros((N, gs)), "bodies": np.zeros((nb, 30))}
    rc = host_lib().alh_ckpt_read(str(path).encode(), C.byref(exp), C.byref(_ckpt_fields(d)))
    if rc:
        raise ALEBGKError(f"alh_ckpt_read({path}): {ERRORS.get(rc, rc)}")
    d.update(step=info["step"], t=info["t"], dim=info["dim"])
    return d


def checkpoint_write(path, step, t, fields, dim=3):
    """Writes a restart file from arrays shaped as checkpoint_read returns."""
    f8 = lambda a: np.ascontiguousarray(a, dtype=np.float64)
    d = {k: f8(fields[k]) for k in ("x", "y", "z", "mom", "mesh", "wall", "grid", "g")}
    d["ints"] = np.ascontiguousarray(fields["ints"], dtype=np.int32)
    d["flags"] = np.ascontiguousarray(fields["flags"], dtype=np.uint8)
    d["bodies"] = f8(fields.get("bodies", np.zeros((0, 30)))).reshape(-1, 30)
    N, gs = d["g"].shape
    info = CkptInfo(dim, N, gs, int(step), float(t), d["bodies"].shape[0])
    rc = host_lib().alh_ckpt_write(str(path).encode(), C.byref(info), C.byref(_ckpt_fields(d)))
    if rc:
        raise ALEBGKError(f"alh_ckpt_write({path}): {ERRORS.get(rc, rc)}")


def copy_calibration(nbytes=256 << 20, reps=2):
    """A copy kernel over nbytes (read once, written once): the reference
    point that scales a profiler's HBM byte counters."""
    L = hip_lib()
    L.ab_copy_calibration.argtypes = [C.c_ulonglong, C.c_int]
    rc = L.ab_copy_calibration(nbytes, reps)
    if rc:
        raise ALEBGKError(f"ab_copy_calibration: {ERRORS.get(rc, rc)}")


class Batch(Solver):
    """Several 2-D solvers' clouds in one context (ab_batch_create): one
    launch per stage steps all of them, each exactly as it would alone.
    Downloads (g(), moments(), ...) give the members' arrays concatenated in
    order; ranges[k] is member k's particle range. scatter() copies g and
    the moments back into the members, which must stay open meanwhile."""

    def __init__(self, solvers):
        self._lib = hip_lib()
        self.members = list(solvers)
        arr = (C.c_void_p * len(self.members))(*[m._ctx.value for m in self.members])
        self._ctx = C.c_void_p()
        rc = self._lib.ab_batch_create(arr, len(self.members), C.byref(self._ctx))
        if rc:
            msgs = "; ".join(m._lib.ab_last_error(m._ctx).decode() for m in self.members)
            self._ctx = C.c_void_p()
            raise ALEBGKError(f"ab_batch_create: {ERRORS.get(rc, rc)}: {msgs}")
        self.cfg = self.members[0].cfg
        self.cloud = None
        self.N = sum(m.N for m in self.members)
        self.dim = 2
        self.gstride = self.members[0].gstride
        ends = np.cumsum([0] + [m.N for m in self.members])
        self.ranges = [(int(a), int(b)) for a, b in zip(ends[:-1], ends[1:])]

    def upload(self, cloud):
        raise ALEBGKError("a batch takes its clouds from its members")

    def save_checkpoint(self, path, step, t):
        raise ALEBGKError("restart files hold one cloud: scatter() the batch and save the members")

    def scatter(self):
        self._check(self._lib.ab_batch_scatter(self._ctx), "ab_batch_scatter")


class Graph:
    """Several solvers' steps captured into one HIP graph (ab_graph_*): each
    solver is an independent branch; launch() replays nsteps steps of all."""

    def __init__(self, solvers, dt, nsteps):
        self._lib = hip_lib()
        self.solvers = list(solvers)
        arr = (C.c_void_p * len(self.solvers))(*[s._ctx.value for s in self.solvers])
        self._g = C.c_void_p()
        rc = self._lib.ab_graph_create(arr, len(self.solvers), C.c_double(dt), int(nsteps),
                                       C.byref(self._g))
        if rc:
            msgs = "; ".join(s._lib.ab_last_error(s._ctx).decode() for s in self.solvers)
            raise ALEBGKError(f"ab_graph_create: {ERRORS.get(rc, rc)}: {msgs}")

    def launch(self):
        rc = self._lib.ab_graph_launch(self._g)
        if rc:
            msgs = "; ".join(s._lib.ab_last_error(s._ctx).decode() for s in self.solvers)
            raise ALEBGKError(f"ab_graph_launch: {ERRORS.get(rc, rc)}: {msgs}")

    def synchronize(self):
        if self._lib.ab_graph_synchronize(self._g):
            raise ALEBGKError("ab_graph_synchronize failed")

    def close(self):
        if self._g:
            self._lib.ab_graph_destroy(self._g)
            self._g = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def check_config(cfg):
    """ab_check_config: the checks ab_create makes, on the host only.
    Returns (status, message); status 0 means ab_create would accept it."""
    buf = C.create_string_buffer(256)
    rc = hip_lib().ab_check_config(C.byref(cfg), buf, C.c_int(len(buf)))
    return int(rc), buf.value.decode()


def velocities(cfg):
    """Node velocities in lin order, shape (nodes, 3)."""
    Nv = cfg.nv
    dv = (cfg.vmax - cfg.vmin) / (Nv - 1)
    v = cfg.vmin + dv * np.arange(Nv)
    if cfg.dim == 3:
        kv, jv, iv = np.meshgrid(np.arange(Nv), np.arange(Nv), np.arange(Nv), indexing="ij")
        return np.stack([v[iv].ravel(), v[jv].ravel(), v[kv].ravel()], axis=1)
    jv, iv = np.meshgrid(np.arange(Nv), np.arange(Nv), indexing="ij")
    return np.stack([v[iv].ravel(), v[jv].ravel(), np.zeros(Nv * Nv)], axis=1)
