"""Benchmark: phase-space updates/s of the ALE-BGK time step on one MI355X.

Workload (the configuration the headline metric is quoted on): cavity3d,
1 um cube, argon 270 K, density 1.0, uniform lid 1 m/s at z = zfront,
Nx = 40 (64,000 particles) x Nv = 10 (1,000 velocity nodes, 6.4e7 values),
quadratic MLS (10-term basis), dt = 1e-11 s, synthetic initial state (the
reference's: gas at rest at the wall temperature).

One step = one full time step of the mixed-precision pipeline: transport
(MLS coefficients factorised once in FP64, FP32 contraction), positivity,
NAD, diffuse wall, moments, the BGK collision towards the discrete Maxwellian
(FP64 Newton per particle, warm-started by an MLP trained on the GPU on the
run's first steps), both mass rescales. The plain FP64 step of the
reference formulation (continuous Maxwellian) is timed for the speedup; the
FP64 formulation with the same discrete collision is run from the same state
for the same number of steps and the JSON line reports the relative moment
error between it and the measured path.

    python bench.py [--steps K] [--warmup W] [--gpus 1]
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0          # MI355X HBM3E, GB/s


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def moment_errors(m, ref, inner):
    import numpy as np
    d = {}
    d["rho"] = float(np.max(np.abs(m[inner, 0] - ref[inner, 0]) / ref[inner, 0]))
    umax = float(np.max(np.abs(ref[inner, 1:4])))
    d["u"] = float(np.max(np.abs(m[inner, 1:4] - ref[inner, 1:4])) / max(umax, 1e-300))
    d["T"] = float(np.max(np.abs(m[inner, 4] - ref[inner, 4]) / ref[inner, 4]))
    return d


def measured_traffic(nx, nv):
    """HBM bytes per transport launch from the committed rocprofv3 counter
    summary of this workload (FETCH_SIZE and WRITE_SIZE, each scaled by what
    the counter reads for a copy kernel of known size), or None."""
    path = os.path.join(ROOT, "profiles", f"k_apply_cavity3d_n{nx}_nv{nv}_pmc.md")
    if not os.path.exists(path):
        return None, None
    for line in open(path):
        if line.startswith("traffic_bytes_per_launch:"):
            return float(line.split(":")[1]), os.path.relpath(path, ROOT)
    return None, None


def cpu_baseline(nx=40, nv=10, order=2):
    """The reference CPU step (oracle/_ref, built from the reference sources)
    on the workload itself: cavity3d at nx^3, Nv=10, quadratic MLS, two
    steps on every host core this process may use (its CPU affinity, or
    OMP_NUM_THREADS where the host sets a smaller share); the steady second
    step is timed (step 1 carries the MOOD re-transport at uniform density).
    Falls back to the oracle port on a 12^3 sample."""
    threads = os.cpu_count() or 1
    try:
        threads = len(os.sched_getaffinity(0))
    except AttributeError:
        pass
    share = os.environ.get("OMP_NUM_THREADS", "")
    if share.isdigit() and int(share) > 0:
        threads = min(threads, int(share))
    exe = os.path.join(ROOT, "oracle", "_ref", "ref_harness_v3")
    if os.path.exists(exe):
        env = dict(os.environ, OMP_NUM_THREADS=str(threads))
        t0 = time.time()
        r = subprocess.run([exe, "cpustep", str(nx), str(nv), str(order), "2"], env=env,
                           stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True,
                           timeout=600)
        if r.returncode == 0:
            out = json.loads(r.stdout.strip().splitlines()[-1])
            sec = out["step_seconds"][-1]
            upd = out["N"] * nv ** 3
            model = ""
            try:
                for ln in open("/proc/cpuinfo"):
                    if ln.startswith("model name"):
                        model = ln.split(":", 1)[1].strip()
                        break
            except OSError:
                pass
            return {"value": upd / sec, "unit": "phase-space updates/s",
                    "cores": out["threads"], "kind": "reference",
                    "sample": (f"reference CPU step (OpenMP, x86-64-v3, {out['threads']} threads"
                               f"{', ' + model if model else ''}) on cavity3d {nx}^3 "
                               f"(N={out['N']}), Nv={nv}, quadratic MLS; steady step 2 "
                               f"timed: {sec:.2f} s; wall {time.time() - t0:.1f} s")}
        log(f"[bench] reference CPU baseline failed: {r.stderr[-400:]}")
    # oracle port: single thread, small sample
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_ctypes as O
    from helpers import setup
    case, cl, ocl, neigh, nn = setup("cavity3d", 12, nv, order=order, perturb=False)
    st = O.State(ocl, cl.g, cl.mom)
    O.step(ocl, st, case.dt, case.tau, case.R, cl.cfg.radius, order)
    t0 = time.perf_counter()
    O.step(ocl, st, case.dt, case.tau, case.R, cl.cfg.radius, order)
    sec = time.perf_counter() - t0
    return {"value": cl.N * nv ** 3 / sec, "unit": "phase-space updates/s", "cores": 1,
            "kind": "port", "sample": f"oracle port, cavity3d 12^3, Nv={nv}, quadratic, step 2"}


def cloud_management(A, np):
    """ab_cloud_manage on the 40^3 cavity (Nv 10, FP64 cloud) with a sphere of
    radius 3.2 dx: the lattice particles inside it removed, 300 Fibonacci
    surface particles added (boundary, body 0), then the body and its surface
    moved by (0.6, 0.3, 0.2) dx and the cloud managed (remove, rebuild, merge
    with minDist 0.5 dx, refill, rebuilds). Synthetic state: the lattice's
    initial condition, the surface rows copied from an interior particle."""
    case = A.configure("cavity3d", nx=40, nv=10, mls_order=2)
    cl = A.build_lattice(case, precision=A.PREC_FP64, mls_order=2)
    xl, xr, yt, yb, zf, zb = case.dom
    dx = (xr - xl) / (case.nx - 1)
    c0 = np.array([xl + 0.52 * (xr - xl), yb + 0.47 * (yt - yb), min(zf, zb) + 0.5 * abs(zb - zf)])
    R = 3.2 * dx
    pos = np.stack([cl.x, cl.y, cl.z], 1)
    keep = (cl.boundary == 1) | (np.sum((pos - c0) ** 2, 1) > R * R)
    n = 300
    k = np.arange(n) + 0.5
    zk = 1.0 - 2.0 * k / n
    rk, ph = np.sqrt(1.0 - zk * zk), k * np.pi * (3.0 - np.sqrt(5.0))
    nrm = np.stack([rk * np.cos(ph), rk * np.sin(ph), zk], 1)
    shift = np.array([0.6, 0.3, 0.2]) * dx
    surf = c0 + shift + R * nrm
    wall = np.zeros((n, 7))
    wall[:, :3], wall[:, 6] = nrm, case.Tw
    interior = int(np.argmax(cl.boundary == 0))
    x = np.concatenate([cl.x[keep], surf[:, 0]])
    y = np.concatenate([cl.y[keep], surf[:, 1]])
    z = np.concatenate([cl.z[keep], surf[:, 2]])
    b = np.concatenate([cl.boundary[keep], np.ones(n, np.uint8)])
    cloud = A.Cloud(cl.cfg, x, y, z, b, np.concatenate([cl.wall[keep], wall]),
                    np.concatenate([cl.mom[keep], np.tile(cl.mom[interior], (n, 1))]),
                    np.concatenate([cl.g[keep], np.tile(cl.g[interior], (n, 1))]))
    body_id = np.concatenate([np.full(int(keep.sum()), -1, np.int32), np.zeros(n, np.int32)])
    sol = A.Solver(cloud, precision=A.PREC_FP64)
    dom6 = (xl, xr, yb, yt, zf, zb)
    sol.synchronize()
    t0 = time.perf_counter()
    src, counts = sol.cloud_manage([A.body(A.BODY_SPHERE, c0 + shift, R)], body_id, dx,
                                   0.5 * dx, dom6)
    ms = (time.perf_counter() - t0) * 1e3
    res = {"ms": ms, "particles_before": int(len(x)), "particles_after": int(sol.N),
           "removed_merged_added": list(counts),
           "note": "remove, neighbour rebuild, merge, refill, rebuilds; one host sync per stage"}
    sol.close()
    log(f"[bench] cloud management 40^3 + sphere: {ms:.2f} ms, counts {counts}")
    return res


def production(A, steps=10):
    """cavity3d at production size: Nx=60 (216,000 particles) x Nv=20 (8,000
    nodes, 1.73e9 values, 6.9 GB in FP32), quadratic MLS, lid 1 m/s, the
    mixed step with the reference's continuous collision: step time and the
    transport contraction's share of the HBM roofline (algorithmic bytes:
    g read once and gt written once, FP32)."""
    case = A.configure("cavity3d", nx=60, nv=20, mls_order=2)
    cl = A.build_lattice(case, precision=A.PREC_MIXED, mls_order=2)
    N = cl.N
    s = A.Solver(cl, precision=A.PREC_MIXED)
    del cl
    shapes, chosen = s.autotune(case.dt, reps=2)
    chosen_name = [name for i, name, _ in shapes if i == chosen][0]
    s.steps(case.dt, 3)
    s.synchronize()
    t0 = time.perf_counter()
    s.steps(case.dt, steps)
    s.synchronize()
    sec = (time.perf_counter() - t0) / steps
    s.profile(True)
    s.steps(case.dt, 5)
    prof = s.profile_read()
    s.profile(False)
    tr_ms, tr_n = prof["transport"]
    tr = tr_ms / max(tr_n, 1) * 1e-3
    upd = N * 20 ** 3
    alg = upd * 2 * 4
    mom = s.moments()
    s.close()
    res = {"ms_per_step": sec * 1e3, "updates_per_s": upd / sec, "particles": N,
           "velocity_nodes": 8000, "transport_ms": tr * 1e3,
           "transport_GBps": alg / tr / 1e9, "transport_hbm_frac": alg / tr / 1e9 / HBM_PEAK_GBS,
           "transport_shape": chosen_name,
           "transport_shapes_ms": {name: ms for _, name, ms in shapes},
           "finite": bool(__import__("numpy").isfinite(mom).all())}
    log(f"[bench] cavity3d 60^3 Nv20 mixed: {sec * 1e3:.3f} ms/step, transport "
        f"{tr * 1e3:.3f} ms ({res['transport_hbm_frac']:.3f} of HBM peak)")
    return res


def cube_case(A, np, steps=10):
    """cavity3d-cube at Nx=50, Nv=10 (density 4.4, parabolic lid peaking at
    10 m/s, a free cube, dt 1.5e-11), quadratic MLS: the FP64 step (a body
    cloud is no lattice) followed by the device moving-body step (Heun loads,
    Newton-Euler, contact, surface wall state, cloud management) every step."""
    cs = A.configure("cavity3d-cube", nx=50, nv=10, mls_order=2)
    cl = A.build_case(cs, mls_order=2)
    s = A.Solver(cl, precision=A.PREC_FP64)
    bodies = cl.bodies
    box = A.box(cs)

    def one():
        s.step(cs.dt)
        return s.body_step(cs.dt, bodies, cl.geo[0], cl.geo[3], box)[1]

    for _ in range(2):
        one()
    s.synchronize()
    t0 = time.perf_counter()
    tot = np.zeros(3, int)
    for _ in range(steps):
        tot += np.array(one())
    s.synchronize()
    sec = (time.perf_counter() - t0) / steps
    t0 = time.perf_counter()
    for _ in range(steps):
        s.step(cs.dt)
    s.synchronize()
    fluid = (time.perf_counter() - t0) / steps
    upd = s.N * 1000
    st = bodies[0].state()
    s.close()
    res = {"ms_per_step": sec * 1e3, "fluid_ms_per_step": fluid * 1e3,
           "body_block_ms": (sec - fluid) * 1e3, "updates_per_s": upd / sec,
           "particles": int(s.N), "removed_merged_added": tot.tolist(),
           "body_omega": st["om"].tolist(), "body_velocity": st["vel"].tolist()}
    log(f"[bench] cavity3d-cube 50^3 Nv10 fp64 + moving body: {sec * 1e3:.2f} ms/step "
        f"(fluid {fluid * 1e3:.2f}, body block {(sec - fluid) * 1e3:.2f})")
    res["learned_warm_start"] = cube_case_dm(A, np)
    return res


def cube_case_dm(A, np, record=4, steps=5):
    """The same case with the discrete-Maxwellian collision (FP64: a body
    cloud is no lattice) and the learned warm start trained on pairs from
    this run's own plain steps: `record` steps recorded (before the first
    body move: a cloud re-layout drops recorded rows), the MLP trained on
    the GPU, then `steps` steps with the moving body, with and without the
    warm start. Reports the step times and the Newton passes per solve."""
    cs = A.configure("cavity3d-cube", nx=50, nv=10, mls_order=2)
    cl = A.build_case(cs, mls_order=2)
    s = A.Solver(cl, precision=A.PREC_FP64, collision=A.COLLISION_DISCRETE)
    bodies = cl.bodies
    box = A.box(cs)
    s.dm_record(record)
    for _ in range(record):
        s.step(cs.dt)
    st_p = s.step(cs.dt, stats=True)
    s.mlp_init(seed=1)
    loss0, loss1 = s.mlp_train(400, batch=4096, lr=2e-3, seed=5)

    def run(on):
        s.mlp_enable(on)
        s.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            s.step(cs.dt)
            s.body_step(cs.dt, bodies, cl.geo[0], cl.geo[3], box)
        s.synchronize()
        st = s.step(cs.dt, stats=True)
        s.body_step(cs.dt, bodies, cl.geo[0], cl.geo[3], box)
        return (time.perf_counter() - t0) / (steps + 1), st

    sec_l, st_l = run(True)
    sec_p, st_q = run(False)
    passes = {"plain": st_p["dm_passes"] / max(st_p["dm_particles"], 1),
              "plain_moving_body": st_q["dm_passes"] / max(st_q["dm_particles"], 1),
              "learned_moving_body": st_l["dm_passes"] / max(st_l["dm_particles"], 1)}
    om = bodies[0].state()["om"].tolist()
    s.close()
    log(f"[bench] cavity3d-cube 50^3 Nv10 fp64 + discrete collision + moving body: "
        f"{sec_l * 1e3:.2f} ms/step with the learned warm start, {sec_p * 1e3:.2f} without; "
        f"Newton passes per solve {passes}")
    return {"ms_per_step_learned": sec_l * 1e3, "ms_per_step_plain": sec_p * 1e3,
            "newton_passes_per_solve": passes, "mlp_loss": [loss0, loss1],
            "recorded_steps": record, "body_omega": om}


def batched_circles(A, np, steps=20):
    """16 independent cavity2d-circle instances (density 11, parabolic lid
    peaking at 10 m/s, a free circle, dt 1.5e-11), Nx drawn from {31, 40, 50}
    and the body scale from U[0.15, 0.3] with a fixed seed, Nv 16, quadratic
    MLS, FP64 (a body cloud is no lattice): every step enqueues the 16 fluid
    steps on their own streams, then runs the 16 moving-body blocks
    (ab_body_step_many, one cloud after another)."""
    import random
    rng = random.Random(20261202)
    specs = [(rng.choice([31, 40, 50]), rng.uniform(0.15, 0.3)) for _ in range(16)]
    sols, bodies, geos, doms, upd = [], [], [], [], 0
    for nx, sc in specs:
        cs = A.configure("cavity2d-circle", nx=nx, nv=16, mls_order=2)
        cl = A.build_case(cs, body_scale=sc, mls_order=2)
        sols.append(A.Solver(cl, precision=A.PREC_FP64))
        bodies.append(cl.bodies)
        geos.append(cl.geo)
        doms.append(A.box(cs))
    dt = cs.dt

    def one():
        for s in sols:
            s.step(dt)
        return A.body_step_many(sols, dt, bodies, geos, doms)

    for _ in range(2):
        one()
    for s in sols:
        s.synchronize()
    tot = np.zeros(3, int)
    t0 = time.perf_counter()
    for _ in range(steps):
        for c in one():
            tot += np.array(c)
    for s in sols:
        s.synchronize()
    sec = (time.perf_counter() - t0) / steps
    upd = sum(s.N for s in sols) * 256
    t0 = time.perf_counter()
    for _ in range(steps):
        for s in sols:
            s.step(dt)
    for s in sols:
        s.synchronize()
    fluid = (time.perf_counter() - t0) / steps
    # the 16 fluid steps as one HIP graph replay per step (16 branches),
    # captured again whenever a body step re-laid a cloud out (the launch is
    # then refused)
    graph = [A.Graph(sols, dt, 1)]
    recaptures = 0

    def one_graph():
        nonlocal recaptures
        try:
            graph[0].launch()
        except A.ALEBGKError:
            graph[0].close()
            graph[0] = A.Graph(sols, dt, 1)
            recaptures += 1
            graph[0].launch()
        return A.body_step_many(sols, dt, bodies, geos, doms)

    one_graph()
    graph[0].synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        one_graph()
    graph[0].synchronize()
    sec_g = (time.perf_counter() - t0) / steps
    t0 = time.perf_counter()
    for _ in range(steps):
        graph[0].launch()
    graph[0].synchronize()
    fluid_g = (time.perf_counter() - t0) / steps
    graph[0].close()
    rot = [float(np.arctan2(b[0].o[1], b[0].o[0])) for b in bodies]
    for s in sols:
        s.close()
    res = {"instances": [{"nx": nx, "body_scale": round(sc, 4)} for nx, sc in specs],
           "nv": 16, "ms_per_step": sec * 1e3, "fluid_ms_per_step": fluid * 1e3,
           "graph_ms_per_step": sec_g * 1e3, "graph_fluid_ms_per_step": fluid_g * 1e3,
           "graph_recaptures": recaptures,
           "updates_per_s": upd / min(sec, sec_g), "removed_merged_added": tot.tolist(),
           "body_rotation_rad_after": rot, "steps": 2 * steps + 3}
    log(f"[bench] 16 cavity2d-circle instances, moving bodies: {sec * 1e3:.3f} ms/step "
        f"(fluid alone {fluid * 1e3:.3f}); fluid steps as one graph: {sec_g * 1e3:.3f} ms/step "
        f"(fluid alone {fluid_g * 1e3:.3f}, {recaptures} recaptures)")
    return res


def components(A):
    """The other rows of the path, measured beside the headline: the neighbour
    rebuild on the 40^3 cloud, cloud management around a moved sphere in it,
    the 2-D Chu step on the reference's own cavity2d case (Nx=31, Nv=16,
    quadratic) with the direct FP64 transport and with the FP64 table
    transport, and 16 independent cavity2d instances (Nx drawn from
    {31, 40, 50} with a fixed seed; plain cavities) replayed in one HIP graph
    against the same instances stepped eagerly, and as one batch context (one
    launch per stage for all instances), eagerly and in a graph, in both
    formulations."""
    import random
    import numpy as np
    out = {}
    try:
        out["cloud_management_40cube_sphere"] = cloud_management(A, np)
    except Exception as e:                          # reported, not hidden
        out["cloud_management_40cube_sphere"] = {"error": str(e)}
    try:
        case = A.configure("cavity3d", nx=40, nv=10, mls_order=2)
        cl = A.build_lattice(case, precision=A.PREC_FP64, mls_order=2)
        s = A.Solver(cl, precision=A.PREC_FP64)
        for _ in range(3):
            s.rebuild_neighbours()
        s.synchronize()
        n = 20
        t0 = time.perf_counter()
        for _ in range(n):
            s.rebuild_neighbours()
        s.synchronize()
        ms = (time.perf_counter() - t0) * 1e3 / n
        out["neighbour_rebuild_40cube"] = {"ms": ms, "particles_per_s": cl.N / ms * 1e3,
                                           "bound": "latency (integer counting sort, LDS-staged search; 6 launches, back to back)"}
        s.close()
        log(f"[bench] neighbour rebuild 40^3: {ms:.3f} ms")
    except Exception as e:                          # reported, not hidden
        out["neighbour_rebuild_40cube"] = {"error": str(e)}
    for tag, prec in (("fp64", A.PREC_FP64), ("fp64_table", A.PREC_FP64_TABLE)):
        key = f"cavity2d_n31_nv16_{tag}"
        try:
            case = A.configure("cavity2d", nx=31, nv=16, mls_order=2)
            cl = A.build_lattice(case, precision=A.PREC_FP64, mls_order=2)
            s = A.Solver(cl, precision=prec)
            s.steps(case.dt, 5)
            s.synchronize()
            n = 50
            t0 = time.perf_counter()
            s.steps(case.dt, n)
            s.synchronize()
            sec = (time.perf_counter() - t0) / n
            upd = cl.N * 16 * 16
            out[key] = {"ms_per_step": sec * 1e3, "updates_per_s": upd / sec,
                        "particles": cl.N, "velocity_nodes": 256}
            s.close()
            log(f"[bench] cavity2d 31^2 Nv16 {tag}: {sec * 1e3:.3f} ms/step")
        except Exception as e:
            out[key] = {"error": str(e)}
    try:
        out["production_cavity3d_n60_nv20_mixed"] = production(A)
    except Exception as e:                          # reported, not hidden
        out["production_cavity3d_n60_nv20_mixed"] = {"error": str(e)}
    try:
        out["batched_cavity2d_circle_16_fp64"] = batched_circles(A, np)
    except Exception as e:                          # reported, not hidden
        out["batched_cavity2d_circle_16_fp64"] = {"error": str(e)}
    try:
        out["cavity3d_cube_n50_nv10_fp64"] = cube_case(A, np)
    except Exception as e:                          # reported, not hidden
        out["cavity3d_cube_n50_nv10_fp64"] = {"error": str(e)}
    rng = random.Random(20261202)
    sizes = [rng.choice([31, 40, 50]) for _ in range(16)]
    for tag, prec in (("fp64", A.PREC_FP64), ("fp64_table", A.PREC_FP64_TABLE)):
        key = f"batched_cavity2d_16_{tag}"
        try:
            sols, upd = [], 0
            for nx in sizes:
                case = A.configure("cavity2d", nx=nx, nv=16, mls_order=2)
                cl = A.build_lattice(case, precision=A.PREC_FP64, mls_order=2)
                sols.append(A.Solver(cl, precision=prec))
                upd += cl.N * 256
            dt = A.configure("cavity2d", nx=31, nv=16, mls_order=2).dt
            per = 10
            graph = A.Graph(sols, dt, per)
            graph.launch()
            graph.synchronize()
            reps = 5
            t0 = time.perf_counter()
            for _ in range(reps):
                graph.launch()
            graph.synchronize()
            sec_g = (time.perf_counter() - t0) / (reps * per)
            graph.close()
            t0 = time.perf_counter()
            for _ in range(reps):
                for sv in sols:
                    sv.steps(dt, per)
            for sv in sols:
                sv.synchronize()
            sec_e = (time.perf_counter() - t0) / (reps * per)
            # the same instances as one batch context: one launch per stage
            batch = A.Batch(sols)
            runs = None
            if prec == A.PREC_FP64_TABLE:          # the apply's run length, timed
                table, chosen = batch.autotune(dt, reps=3)
                runs = {"apply_ms": {name: ms for _, name, ms in table},
                        "chosen": table[chosen][1]}
            batch.steps(dt, per)
            batch.synchronize()
            t0 = time.perf_counter()
            batch.steps(dt, reps * per)
            batch.synchronize()
            sec_b = (time.perf_counter() - t0) / (reps * per)
            bgraph = A.Graph([batch], dt, per)
            bgraph.launch()
            bgraph.synchronize()
            t0 = time.perf_counter()
            for _ in range(reps):
                bgraph.launch()
            bgraph.synchronize()
            sec_bg = (time.perf_counter() - t0) / (reps * per)
            bgraph.close()
            batch.close()
            out[key] = {"nx": sizes, "nv": 16, "graph_ms_per_step": sec_g * 1e3,
                        "eager_ms_per_step": sec_e * 1e3,
                        "graph_updates_per_s": upd / sec_g,
                        "eager_updates_per_s": upd / sec_e,
                        "batch_ms_per_step": sec_b * 1e3,
                        "batch_graph_ms_per_step": sec_bg * 1e3,
                        "batch_graph_updates_per_s": upd / sec_bg}
            if runs:
                out[key]["batch_run_lengths"] = runs
            for sv in sols:
                sv.close()
            log(f"[bench] 16 cavity2d instances {tag}: graph {sec_g * 1e3:.3f} ms/step, "
                f"eager {sec_e * 1e3:.3f} ms/step, batch {sec_b * 1e3:.3f} ms/step, "
                f"batch in a graph {sec_bg * 1e3:.3f} ms/step")
        except Exception as e:
            out[key] = {"error": str(e)}
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--nx", type=int, default=40)
    ap.add_argument("--nv", type=int, default=10)
    ap.add_argument("--record", type=int, default=8,
                    help="warmup steps whose discrete-Maxwellian solves train the MLP")
    ap.add_argument("--train-iters", type=int, default=600)
    ap.add_argument("--fp64-steps", type=int, default=None,
                    help="FP64 steps for the comparison (default: warmup + steps)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--counters", action="store_true",
                    help="for rocprofv3 --pmc: the measured step only (warmup, training, "
                         "autotune, --steps steps), then a 256 MiB copy kernel that calibrates "
                         "the byte counters; prints no JSON")
    ap.add_argument("--batch2d", action="store_true",
                    help="for rocprofv3: only the 16 cavity2d instances (Nx from {31, 40, 50}, "
                         "Nv 16, quadratic) as one batch context with the 2-D table transport, "
                         "its run length autotuned, "
                         "5 warmup and --steps steps, then a 256 MiB copy kernel that "
                         "calibrates the byte counters; prints the step time, no JSON")
    args = ap.parse_args()
    if args.batch2d:
        import random
        import alebgk4paper_amd as A
        rng = random.Random(20261202)
        sols = []
        for nx in [rng.choice([31, 40, 50]) for _ in range(16)]:
            case = A.configure("cavity2d", nx=nx, nv=16, mls_order=2)
            cl = A.build_lattice(case, precision=A.PREC_FP64, mls_order=2)
            sols.append(A.Solver(cl, precision=A.PREC_FP64_TABLE))
        batch = A.Batch(sols)
        dt = A.configure("cavity2d", nx=31, nv=16, mls_order=2).dt
        table, chosen = batch.autotune(dt, reps=3)
        log(f"[bench] 2-D apply run lengths: " +
            ", ".join(f"{name} {ms * 1e3:.1f} us" for _, name, ms in table) +
            f"; kept {table[chosen][1]}")
        batch.steps(dt, 5)
        batch.synchronize()
        t0 = time.perf_counter()
        batch.steps(dt, args.steps)
        batch.synchronize()
        log(f"[bench] 16 cavity2d instances as one batch (table transport): "
            f"{(time.perf_counter() - t0) / args.steps * 1e3:.3f} ms/step, N={batch.N}")
        A.copy_calibration(256 << 20, 2)     # scales a profiler's byte counters
        return
    if args.gpus != 1:
        print(f"bench.py runs on one GPU: this build has no multi-GPU path "
              f"(--gpus {args.gpus} requested)", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import alebgk4paper_amd as A
    DISC = A.COLLISION_DISCRETE

    case = A.configure("cavity3d", nx=args.nx, nv=args.nv, mls_order=2)
    cloud = A.build_lattice(case, precision=A.PREC_MIXED, mls_order=2)
    N, nodes = cloud.N, args.nv ** 3
    updates = N * nodes
    dt = case.dt
    rec = min(args.record, args.warmup)
    log(f"[bench] cavity3d {args.nx}^3 Nv={args.nv}: N={N}, {updates:.3e} updates/step")

    # ---- the measured step: mixed precision, discrete-Maxwellian collision,
    # Newton warm-started by the MLP trained on this run's own solves -------
    mix = A.Solver(cloud, precision=A.PREC_MIXED, collision=DISC)
    mix.steps(dt, args.warmup - rec)
    mix.dm_record(rec)
    mix.steps(dt, rec)
    mix.synchronize()
    t0 = time.perf_counter()
    mix.mlp_init(seed=1)
    loss0, loss1 = mix.mlp_train(args.train_iters, batch=4096, lr=2e-3, seed=5)
    mix.mlp_enable(True)
    train_s = time.perf_counter() - t0
    shapes, chosen = mix.autotune(dt, reps=3)
    shape_ms = {name: ms for _, name, ms in shapes}
    chosen_name = [name for i, name, _ in shapes if i == chosen][0]
    log(f"[bench] transport shapes (ms): {shape_ms} -> {chosen_name}")
    log(f"[bench] MLP: {args.train_iters} Adam steps on {rec * N} rows in {train_s:.2f} s, "
        f"loss {loss0:.3e} -> {loss1:.3e}")
    if args.counters:
        mix.steps(dt, args.steps)
        mix.synchronize()
        A.copy_calibration(256 << 20, 2)
        log("[bench] counters pass done")
        return
    t0 = time.perf_counter()
    mix.steps(dt, args.steps)
    mix.synchronize()
    t1 = time.perf_counter()
    sec = t1 - t0
    value = updates * args.steps / sec
    log(f"[bench] mixed+DM+MLP: {sec * 1e3 / args.steps:.4f} ms/step, {value:.4e} updates/s")
    m_mix = mix.moments()                 # after warmup + steps steps: the accuracy check
    # stage times from HIP events around each stage, in a run of their own
    # (the events add their own time to the step)
    nprof = max(1, min(args.steps, 50))
    mix.profile(True)
    mix.steps(dt, nprof)
    mix.synchronize()
    prof = mix.profile_read()
    mix.profile(False)
    log(f"[bench] stage times over {nprof} steps (ms total, launches): {prof}")
    tr_ms, tr_n = prof["transport"]
    tr_avg_s = tr_ms / max(tr_n, 1) * 1e-3
    alg_bytes = updates * 2 * 4           # read d once + write gt once, FP32
    achieved = alg_bytes / tr_avg_s / 1e9
    st_l = mix.step(dt, stats=True)       # one more step: Newton passes with the warm start

    # ---- the same without the warm start (plain Newton) --------------------
    pln = A.Solver(cloud, precision=A.PREC_MIXED, collision=DISC)
    pln.steps(dt, args.warmup)
    pln.set_transport_shape(chosen)
    pln.synchronize()
    t0 = time.perf_counter()
    pln.steps(dt, args.steps)
    pln.synchronize()
    sec_plain = time.perf_counter() - t0
    st_p = pln.step(dt, stats=True)
    value_plain = updates * args.steps / sec_plain
    log(f"[bench] mixed+DM (no warm start): {sec_plain * 1e3 / args.steps:.4f} ms/step")
    passes = {"plain": st_p["dm_passes"] / max(st_p["dm_particles"], 1),
              "learned": st_l["dm_passes"] / max(st_l["dm_particles"], 1)}
    log(f"[bench] Newton grid passes per solve: {passes}")
    pln.close()

    # ---- the mixed step with the reference's continuous collision -----------
    mcont = A.Solver(cloud, precision=A.PREC_MIXED)
    mcont.set_transport_shape(chosen)
    mcont.steps(dt, args.warmup)
    mcont.synchronize()
    t0 = time.perf_counter()
    mcont.steps(dt, args.steps)
    mcont.synchronize()
    sec_mc = (time.perf_counter() - t0) / args.steps
    log(f"[bench] mixed, continuous collision: {sec_mc * 1e3:.4f} ms/step")
    mcont.close()

    # ---- the reference formulation in FP64 ----------------------------------
    n64 = args.fp64_steps if args.fp64_steps is not None else args.warmup + args.steps
    ref = A.Solver(cloud, precision=A.PREC_FP64)      # the plain FP64 step (continuous BGK)
    ref.step(dt)
    ref.synchronize()
    t0 = time.perf_counter()
    ref.steps(dt, n64 - 1)
    ref.synchronize()
    sec64 = (time.perf_counter() - t0) / max(n64 - 1, 1)
    value64 = updates / sec64
    log(f"[bench] fp64: {sec64 * 1e3:.3f} ms/step, {value64:.4e} updates/s")
    ref.close()

    # ---- FP64 with the same algorithm (tables, discrete collision): the
    # comparator of the accuracy check, and the speed-up of precision alone --
    tab = A.Solver(cloud, precision=A.PREC_FP64_TABLE, collision=DISC)
    tab.steps(dt, args.warmup)
    tab.synchronize()
    t0 = time.perf_counter()
    tab.steps(dt, args.steps)
    tab.synchronize()
    sec_t64 = (time.perf_counter() - t0) / args.steps
    log(f"[bench] fp64 tables + DM: {sec_t64 * 1e3:.4f} ms/step")
    # the mixed path's moments were taken after warmup + steps steps
    err = moment_errors(m_mix, tab.moments(), cloud.boundary == 0)
    within = bool(max(err.values()) <= 1e-5)
    log(f"[bench] mixed+DM+MLP vs FP64 tables+DM moment errors after "
        f"{args.warmup + args.steps} steps: {err} (within 1e-5: {within})")
    tab.close()

    comps = components(A)
    traffic, traffic_src = measured_traffic(args.nx, args.nv)
    cpu = None if args.no_cpu else cpu_baseline(nv=args.nv)
    if cpu:
        log(f"[bench] cpu baseline: {cpu}")

    line = {
        "metric": "phase-space updates/s, cavity3d Nx40 Nv10: mixed vs FP64 GPU @1e-5 moment err",
        "value": value,
        "unit": "phase-space updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": sec * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic: the reference's initial state (gas at rest, 270 K) on the cavity3d lattice",
        "config": {"workload": f"cavity3d Nx={args.nx} Nv={args.nv} quadratic MLS, dt=1e-11, "
                               f"mixed precision (FP32 transport and collision; FP64 MLS "
                               f"factorisation, conservation, wall scaling and discrete-Maxwellian "
                               f"Newton), MLP warm start trained on the run's first {rec} steps",
                   "particles": N, "velocity_nodes": nodes, "updates_per_step": updates},
        "fp64_value": value64,
        "fp64_ms_per_step": sec64 * 1e3,
        "speedup_vs_fp64": value / value64,
        "fp64_table_dm_ms_per_step": sec_t64 * 1e3,
        "speedup_vs_fp64_same_algorithm": sec_t64 / (sec / args.steps),
        "mixed_continuous_ms_per_step": sec_mc * 1e3,
        "moment_err_vs_fp64": err,
        "moment_err_reference": "FP64 tables with the discrete collision, same steps",
        "within_1e-5": within,
        "without_warm_start": {"value": value_plain, "ms_per_step": sec_plain * 1e3 / args.steps},
        "newton_passes_per_solve": passes,
        "mlp": {"train_seconds": train_s, "adam_steps": args.train_iters, "batch": 4096,
                "rows": rec * N, "loss_first": loss0, "loss_last": loss1},
        "autotune": {"transport_shapes_ms": shape_ms, "chosen": chosen_name},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "traffic_source": traffic_src,
                     "kernel": "k_apply_mix (MLS transport contraction, autotuned shape)",
                     "kernel_avg_ms": tr_avg_s * 1e3,
                     "algorithmic_bytes_per_launch": alg_bytes},
        "cpu_baseline": cpu,
        "components": comps,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
