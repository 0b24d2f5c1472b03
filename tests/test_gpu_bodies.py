"""GPU: moving rigid bodies. The device body step (ab_body_step: Heun-coupled
surface loads, Newton-Euler in FP64, box contact, surface wall state, cloud
management) after the reference's CPU fluid step (AB_LIMITER_MOOD), against
the reference's own functions run for 40 steps by the harness (fixtures
bodysteps_*: runSolver's CPU loop with moveRigidBodies and the cloud
management, Solver/Solver.cpp:632-714).

Tolerances:
  body state every step         velocity, spin, loads 1e-8 relative to each
                                  vector's size (the loads are reduced in
                                  another order, and 40 steps of the fluid in
                                  FP64 on the GPU); centre 1e-10, orientation
                                  1e-9
  body rotation angle after 40 steps    within 0.1 % (north star), measured
                                  ~1e-10
  cloud at the end: counts, order, flags and body ids identical; positions
                                  to 1e-16 m (1e-9 of dx); g and moments 1e-8
"""
from __future__ import annotations

import numpy as np
import pytest

import alebgk4paper_amd as A
from conftest import load_fixture
from helpers import rel_err

pytestmark = pytest.mark.gpu

CASES = {"bodysteps_cube_n12_v4": ("cavity3d-cube", 12, 4),
         "bodysteps_circle_n24_v6": ("cavity2d-circle", 24, 6)}


def _say(msg):
    print(msg, flush=True)


def _start(name):
    """The fixture's starting state on a cloud configured by the host case
    builder (whose body-case clouds are bit-exact to the reference's)."""
    fx = load_fixture(name)
    case_name, nx, nv = CASES[name]
    scale, a, order, nsteps = fx["params"]
    cs = A.configure(case_name, nx=nx, nv=nv, mls_order=int(order))
    cl = A.build_case(cs, body_scale=float(scale), mls_order=int(order))
    assert cl.N == fx.meta["N"]
    for k in ("x", "y", "z"):
        assert np.array_equal(getattr(cl, k), fx[k])
    assert np.array_equal(cl.body_id, fx["bodyid"])
    cl.g = np.ascontiguousarray(fx["g0"])
    cl.mom = np.ascontiguousarray(fx["mom0"])
    cl.wall = np.ascontiguousarray(fx["wall0"])
    return fx, cs, cl, int(nsteps)


def _state_row(b):
    """The fixture's body record layout (oracle/ref_harness.cpp bodyRecord)."""
    r = np.zeros(36)
    r[5:8] = b.c[:]
    r[8:10] = b.o[:]
    r[10:19] = b.rm[:]
    r[19:22] = b.vel[:]
    r[22] = b.angular_vel
    r[23:26] = b.om[:]
    r[26:29] = b.force[:]
    r[29:32] = b.tq[:]
    return r


def _vec_err(a, b):
    den = max(np.abs(b).max(), 1e-300)
    return float(np.abs(a - b).max() / den)


def _angle(r, dim):
    if dim == 3:
        tr = r[10] + r[14] + r[18]
        return float(np.arccos(np.clip((tr - 1.0) / 2.0, -1.0, 1.0)))
    return float(np.arctan2(r[9], r[8]))


@pytest.mark.parametrize("name", list(CASES))
def test_body_steps_match_reference(name):
    fx, cs, cl, nsteps = _start(name)
    _say(f"[bodies] {name}: {nsteps} CPU-semantics steps with a free body")
    dim = cl.cfg.dim
    sol = A.Solver(cl, precision=A.PREC_FP64, limiter=A.LIMITER_MOOD)
    bodies = cl.bodies
    dx, min_dist = cl.geo[0], cl.geo[3]
    hist, counts = fx["body_hist"], fx["counts"]
    worst = {}
    for st in range(nsteps):
        sol.step(cs.dt)
        loads, cnt = sol.body_step(cs.dt, bodies, dx, min_dist, A.box(cs))
        assert (sol.N, *cnt) == tuple(counts[st]), st
        got, want = _state_row(bodies[0]), hist[st]
        for key, sl in (("c", slice(5, 8)), ("vel", slice(19, 22)), ("om", slice(22, 26)),
                        ("force", slice(26, 29)), ("tq", slice(29, 32)), ("orient", slice(8, 19))):
            e = _vec_err(got[sl], want[sl])
            worst[key] = max(worst.get(key, 0.0), e)
    _say(f"  worst relative errors of the body state over {nsteps} steps: {worst}")
    assert worst["c"] < 1e-10, worst
    for key in ("vel", "om", "force", "tq"):
        assert worst[key] < 1e-8, (key, worst)
    assert worst["orient"] < 1e-9, worst
    # the rotation the body accumulated, within 0.1 % of the reference's
    th_ref = _angle(hist[-1], dim) - _angle(fx["body0"], dim)
    th = _angle(_state_row(bodies[0]), dim) - _angle(fx["body0"], dim)
    _say(f"  rotation after {nsteps} steps: {th:.9e} rad (reference {th_ref:.9e})")
    assert abs(th_ref) > 1e-6
    assert abs(th - th_ref) <= 1e-3 * abs(th_ref)
    # the cloud at the end
    p = sol.particles()
    assert np.array_equal(p["boundary"], fx["boundary1"])
    assert np.array_equal(sol.body_ids(), fx["bodyid1"])
    for k in ("x", "y", "z"):
        assert np.abs(p[k] - fx[k + "1"]).max() <= 1e-16, k
    w1 = fx["wall1"]
    assert rel_err(p["wall"][:, :3], w1[:, :3]) < 1e-12
    assert rel_err(p["wall"][:, 3:6], w1[:, 3:6]) < 1e-8
    assert rel_err(sol.g(), fx["g1"]) < 1e-8
    mom = sol.moments()
    ref = fx["mom1"]
    cols = [0, 1, 2, 3, 4] if dim == 3 else [0, 1, 2, 4]
    for c in cols:
        assert rel_err(mom[:, c], ref[:, c], floor=1e-300) < 1e-8, c


def test_body_loads_on_device_match_host_sum():
    """ab_body_loads reduces on the device: against the same loads summed
    on the host from the downloaded g and wall state (the reference's
    computeBodyLoads formula, BodyDynamics.cpp:41-94), for the current wall
    velocities and for given ones."""
    fx, cs, cl, nsteps = _start("bodysteps_cube_n12_v4")
    # the streams load the cube symmetrically (the net force nearly cancels):
    # a random factor per value gives loads of their own size
    rng = np.random.default_rng(5)
    cl.g = np.ascontiguousarray(cl.g * (1.0 + 0.3 * rng.uniform(-1, 1, size=cl.g.shape)))
    sol = A.Solver(cl, precision=A.PREC_FP64)
    surf = np.nonzero((cl.boundary == 1) & (cl.body_id >= 0))[0].astype(np.int32)
    b = cl.bodies[0]
    area = 6.0 * b.size * b.size / b.npts
    v = A.velocities(cl.cfg)
    dv3 = ((cl.cfg.vmax - cl.cfg.vmin) / (cl.cfg.nv - 1)) ** 3

    def host(wu):
        F = np.zeros(6)
        for k, i in enumerate(surf):
            c = v - wu[k]
            f = cl.g[i] * dv3
            P = np.einsum("n,na,nb->ab", f, c, c)
            t = -P @ cl.wall[i, :3]
            r = np.array([cl.x[i], cl.y[i], cl.z[i]]) - np.array(b.c[:])
            F[:3] += area * t
            F[3:] += area * np.cross(r, t)
        return F

    wu = rng.normal(scale=5.0, size=(len(surf), 3))
    got0 = sol.body_loads(surf, np.zeros(len(surf), np.int32), [b.c[:]], [area])[0]
    got1 = sol.body_loads(surf, np.zeros(len(surf), np.int32), [b.c[:]], [area], wu=wu)[0]
    for got, want in ((got0, host(cl.wall[surf, 3:6])), (got1, host(wu))):
        _say(f"  device {got}\n  host   {want}")
        assert np.abs(got[:3] - want[:3]).max() <= 1e-11 * np.abs(want[:3]).max()
        assert np.abs(got[3:] - want[3:]).max() <= 1e-11 * np.abs(want[3:]).max()


def test_cube_case_gpu_semantics_runs():
    """cavity3d-cube (density 4.4, parabolic lid, free cube, dt 1.5e-11) at
    a reduced lattice with the GPU step: the body stays in the box, the field
    stays finite and positive, and the cloud management keeps the surface."""
    cs = A.configure("cavity3d-cube", nx=20, nv=6, mls_order=2)
    cl = A.build_case(cs, mls_order=2)
    _say(f"[bodies] cavity3d-cube 20^3 nv6: N={cl.N}, 30 GPU-semantics steps")
    sol = A.Solver(cl, precision=A.PREC_FP64)
    bodies = cl.bodies
    nsurf0 = int(((cl.boundary == 1) & (cl.body_id >= 0)).sum())
    tot = np.zeros(3, int)
    for _ in range(30):
        sol.step(cs.dt)
        _, cnt = sol.body_step(cs.dt, bodies, cl.geo[0], cl.geo[3], A.box(cs))
        tot += np.array(cnt)
    mom = sol.moments()
    assert np.all(np.isfinite(mom)) and np.all(mom[:, 0] > 0)
    s = bodies[0].state()
    assert np.all(np.isfinite(s["vel"])) and np.all(np.isfinite(s["om"]))
    assert np.all(s["c"] > 0) and np.all(s["c"] < 1e-6)
    bid = sol.body_ids()
    assert int(((sol.particles()["boundary"] == 1) & (bid >= 0)).sum()) == nsurf0
    _say(f"  body v={s['vel']}, w={s['om']}, cloud counts {tot}")


def test_many_instances_match_each_alone():
    """Several 2-D cavities with a free circle (uneven lattices and body
    sizes) stepped together -- the fluid steps enqueued on every context's
    stream, or replayed as one HIP graph, then ab_body_step_many -- end
    bit-identical to each instance stepped alone: the same body rotation,
    state and cloud."""
    specs = [(24, 0.25), (28, 0.3), (32, 0.2), (24, 0.15)]
    _say(f"[bodies] {len(specs)} cavity2d-circle instances (nx, body scale) {specs}, nv 8, 12 steps")

    def make():
        out = []
        for nx, sc in specs:
            cs = A.configure("cavity2d-circle", nx=nx, nv=8, mls_order=2)
            cl = A.build_case(cs, body_scale=sc * 4.0, mls_order=2)
            out.append((cs, cl, A.Solver(cl, precision=A.PREC_FP64), cl.bodies))
        return out

    alone, many, graphed = make(), make(), make()
    dt = alone[0][0].dt
    graph = A.Graph([m[2] for m in graphed], dt, 1)
    for _ in range(12):
        for cs, cl, s, b in alone:
            s.step(dt)
            s.body_step(dt, b, cl.geo[0], cl.geo[3], A.box(cs))
        for cs, cl, s, b in many:
            s.step(dt)
        A.body_step_many([m[2] for m in many], dt, [m[3] for m in many],
                         [m[1].geo for m in many], [A.box(m[0]) for m in many])
        # the fluid steps as one graph of four branches, captured again when
        # a body step has re-laid a cloud out
        try:
            graph.launch()
        except A.ALEBGKError:
            graph.close()
            graph = A.Graph([m[2] for m in graphed], dt, 1)
            graph.launch()
        A.body_step_many([m[2] for m in graphed], dt, [m[3] for m in graphed],
                         [m[1].geo for m in graphed], [A.box(m[0]) for m in graphed])
    graph.close()
    for (csa, cla, sa, ba), (csm, clm, sm, bm), (csg, clg, sg, bg) in zip(alone, many, graphed):
        assert sa.N == sm.N == sg.N
        for s2, b2 in ((sm, bm), (sg, bg)):
            assert np.array_equal(_state_row(ba[0]), _state_row(b2[0]))
            assert np.array_equal(sa.g(), s2.g())
            assert np.array_equal(sa.body_ids(), s2.body_ids())
    rot = [float(np.arctan2(m[3][0].o[1], m[3][0].o[0])) for m in many]
    _say(f"  rotations after 12 steps (rad): {rot}")
    assert all(abs(r) > 0 for r in rot)


def test_cube_case_learned_warm_start_with_moving_body():
    """The cube case with the discrete-Maxwellian collision (FP64) and the
    learned warm start trained on the run's own plain steps, then stepped
    with the moving body: the warm start cuts the Newton passes, every solve
    converges, and the body still turns."""
    cs = A.configure("cavity3d-cube", nx=16, nv=6, mls_order=2)
    cl = A.build_case(cs, mls_order=2)
    _say(f"[bodies] cavity3d-cube 16^3 nv6 FP64 + discrete collision, N={cl.N}")
    sol = A.Solver(cl, precision=A.PREC_FP64, collision=A.COLLISION_DISCRETE)
    bodies = cl.bodies
    sol.dm_record(3)
    for _ in range(3):
        sol.step(cs.dt)
    plain = sol.step(cs.dt, stats=True)
    sol.mlp_init(seed=1)
    l0, l1 = sol.mlp_train(300, batch=1024, lr=2e-3, seed=5)
    assert l1 < l0
    sol.mlp_enable(True)
    for _ in range(4):
        sol.step(cs.dt)
        sol.body_step(cs.dt, bodies, cl.geo[0], cl.geo[3], A.box(cs))
    learned = sol.step(cs.dt, stats=True)
    p0 = plain["dm_passes"] / plain["dm_particles"]
    p1 = learned["dm_passes"] / learned["dm_particles"]
    _say(f"  Newton passes per solve: plain {p0:.3f}, learned {p1:.3f}; loss {l0:.3e} -> {l1:.3e}")
    assert learned["dm_unconverged"] == 0
    assert p1 < p0
    mom = sol.moments()
    assert np.all(np.isfinite(mom)) and np.all(mom[:, 0] > 0)
    assert abs(bodies[0].state()["om"][2]) > 0
