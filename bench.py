#!/usr/bin/env python3
"""Benchmark of the PFEM per-time-step hot path on one MI355X.

Headline workload: the serving stream — 512 independent cases (seeded), 60 % 2D / 40 % 3D,
node counts log-uniform in [2k, 300k], half incompressible PSPG dam-break columns, half
weakly compressible sloshing tanks; each request is one time step of one case (a Picard
step with Ax_f minRes 1e-6, maxIter 10 and linear solves to a true relative residual of
1e-8; or a CDS_dpdt step followed by computeNextDT).  Two paths serve the same requests:

* plain FP64: one request at a time through the direct C-ABI calls (FP64 BiCGSTAB with FP64
  multigrid or block Jacobi, FP64 WC state) — the drop-in replacement of the reference's
  per-case step;
* batched mixed: the continuous-batching server (requests join the running batch at every
  engine tick, admitted by step-workspace footprint), PFEM_PREC_MIXED (FP64 BiCGSTAB; the
  multigrid of 3D cases in FP32, from 5,000 nodes), FP32 WC state with FP64 sums.

Each path runs the stream twice: saturated (every request queued at t0: throughput, the
headline steps/s) and with seeded Poisson arrivals at 200 requests/s (p50/p99 latency from
arrival to result).  States are reset to the initial condition before every run, so both
paths advance the same states by one step.

Also measured: the 3D PSPG dam break (262,701 nodes, 1.5M tets) as a plain FP64 drop-in,
with the BiCGSTAB SpMV timed live with HIP events (the roofline of the dominant kernel),
and the CPU oracle on a bounded sample of the stream (the CPU baseline).

Prints ONE JSON line.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
# the CPU baseline's OpenMP threads stay on their cores (read when the oracle library loads)
os.environ.setdefault("OMP_PROC_BIND", "close")
os.environ.setdefault("OMP_PLACES", "cores")

METRIC = "steps/s + p50/p99 step latency, batched mixed vs plain FP64 GPU, relres≤1e-8"
HBM_PEAK_GBPS = 8000.0
K0, K0P, RHO_STAR = 2.2e9, 7.0, 1000.0
RATE = 200.0
WC_SC, WC_MAXDT = 0.5, 1e-3


def log(msg):
    print(msg, file=sys.stderr, flush=True)


# ------------------------------------------------------------------------------------------
# the stream
# ------------------------------------------------------------------------------------------
def build_stream(n, seed=2024):
    """Seeded request list: kind, dim, target node count and Poisson arrival offset (s)."""
    rng = np.random.default_rng(seed)
    dims = np.where(rng.random(n) < 0.6, 2, 3)
    nodes = np.exp(rng.uniform(math.log(2000), math.log(300000), n)).astype(int)
    kinds = np.where(rng.random(n) < 0.5, "PSPG", "WC")
    gaps = rng.exponential(1.0 / RATE, n)
    arrivals = np.cumsum(gaps) - gaps[0]
    return [dict(i=i, kind=str(kinds[i]), dim=int(dims[i]), target=int(nodes[i]), arrival=float(arrivals[i]))
            for i in range(n)]


def make_mesh(req):
    from pfem_amd import meshgen as mg

    seed = 1000 + req["i"]
    if req["kind"] == "PSPG":
        if req["dim"] == 2:
            return mg.column_2d_nodes(req["target"], h=0.01, seed=seed)
        return mg.column_3d_nodes(req["target"], h=0.01, seed=seed)
    return mg.tank_nodes(req["dim"], req["target"], h=0.01, seed=seed)


def initial_state(req, mesh):
    from pfem_amd import meshgen as mg

    if req["kind"] == "PSPG":
        return mg.pspg_hydrostatic(mesh)
    v, p, rho, a = mg.wc_tait_hydrostatic(mesh, K0=K0, K0p=K0P, rho_star=RHO_STAR)
    return np.concatenate([v.reshape(-1), p, rho, a.reshape(-1)])


def make_case(ctx, req, mesh, variant, batched_precision="mixed"):
    import pfem_amd

    if req["kind"] == "PSPG":
        c = pfem_amd.Case(ctx, req["dim"], kind="PSPG", rho=1000.0, mu=1e-3, residual="Ax_f", min_res=1e-6,
                          max_iter=10, precision=batched_precision if variant == "batched" else "fp64",
                          krylov_rtol=1e-8,
                          krylov_max_iter=20000, inner_rtol=1e-3, inner_max_iter=300)
    else:
        c = pfem_amd.Case(ctx, req["dim"], kind="WC", mu=1e-3, K0=K0, K0p=K0P, rho_star=RHO_STAR,
                          wc_state="fp32" if variant == "batched" else "fp64")
    c.set_mesh(mesh)
    return c


def pct(a, q):
    return float(np.percentile(np.asarray(a), q)) if len(a) else None


# ------------------------------------------------------------------------------------------
# the two paths
# ------------------------------------------------------------------------------------------
def run_plain(stream, cases, poisson):
    """One request at a time through the direct calls, in arrival order."""
    import pfem_amd

    lat, fails, relres, kit = [], 0, 0.0, 0
    t0 = pfem_amd.now_ns()
    for req in stream:
        c = cases[req["i"]]
        arr = t0 + int(req["arrival"] * 1e9) if poisson else t0
        while poisson and pfem_amd.now_ns() < arr:
            time.sleep(max(0.0, (arr - pfem_amd.now_ns()) * 1e-9 - 2e-4))
        if req["kind"] == "PSPG":
            rc, st = c.pspg_step(req["dt"])
            fails += rc != 0
            relres = max(relres, st["max_relres"])
            kit += st["krylov_iters"]
        else:
            c.wc_step(req["dt"])
            req["dt"] = c.wc_next_dt(WC_SC, WC_MAXDT)  # the next step's dt, as the time loop uses it
        lat.append((pfem_amd.now_ns() - arr) * 1e-6)
    total = (pfem_amd.now_ns() - t0) * 1e-9
    return dict(steps_per_s=len(stream) / total, seconds=total, p50_ms=pct(lat, 50), p99_ms=pct(lat, 99),
                failed=int(fails), max_relres=relres, krylov_iters=int(kit))


def run_batched(ctx, stream, cases, poisson, budget, slots):
    import pfem_amd

    srv = pfem_amd.Server(ctx, hbm_budget=budget, max_inflight=slots, wc_security_coeff=WC_SC, wc_max_dt=WC_MAXDT)
    t0 = pfem_amd.now_ns() + 2_000_000
    for req in stream:
        arr = t0 + int(req["arrival"] * 1e9) if poisson else t0
        srv.submit(cases[req["i"]], req["dt"], tag=req["i"], arrival_ns=arr)
    res = []
    deadline = time.time() + 600
    while len(res) < len(stream):
        if time.time() > deadline:
            raise SystemExit("batched run did not finish")
        res += srv.poll(timeout_ms=100)
    st = srv.stats()
    srv.close()
    by_tag = {r["i"]: r for r in stream}
    for r in res:
        if r["kind"] == 1:
            by_tag[r["tag"]]["dt"] = r["next_dt"]
    t_end = max(r["t_done_ns"] for r in res)
    total = (t_end - t0) * 1e-9
    lat = [(r["t_done_ns"] - r["t_arrival_ns"]) * 1e-6 for r in res]
    pspg = [r for r in res if r["kind"] == 0]
    return dict(steps_per_s=len(stream) / total, seconds=total, p50_ms=pct(lat, 50), p99_ms=pct(lat, 99),
                failed=int(sum(r["status"] != 0 for r in res)),
                max_relres=max([r["stats"]["max_relres"] for r in pspg], default=0.0),
                krylov_iters=int(sum(r["stats"]["krylov_iters"] for r in pspg)), ticks=st["ticks"],
                peak_inflight=st["peak_inflight"], arena_bytes=st["arena_bytes"], arena_peak=st["arena_peak"])


def run_batched_passes(ctx, stream, cases, passes, budget, slots):
    """`passes` saturated passes of the stream through one server: every request of a pass is
    submitted at the pass's start, the next pass starts when the last result is in.  Each
    pass advances every case by one time step (WC cases with the dt computeNextDT gave)."""
    import pfem_amd

    srv = pfem_amd.Server(ctx, hbm_budget=budget, max_inflight=slots, wc_security_coeff=WC_SC, wc_max_dt=WC_MAXDT)
    by_tag = {r["i"]: r for r in stream}
    pass_s, fails, relres, kit1 = [], 0, 0.0, 0
    t_all = pfem_amd.now_ns()
    for k in range(passes):
        t0 = pfem_amd.now_ns()
        for req in stream:
            srv.submit(cases[req["i"]], req["dt"], tag=req["i"], arrival_ns=t0)
        res = []
        deadline = time.time() + 600
        while len(res) < len(stream):
            if time.time() > deadline:
                raise SystemExit("batched run did not finish")
            res += srv.poll(timeout_ms=100)
        t1 = max(r["t_done_ns"] for r in res)
        pass_s.append((t1 - t0) * 1e-9)
        for r in res:
            fails += r["status"] != 0
            if r["kind"] == 0:
                relres = max(relres, r["stats"]["max_relres"])
                if k == 0:
                    kit1 += r["stats"]["krylov_iters"]
            else:
                by_tag[r["tag"]]["dt"] = r["next_dt"]
    total = (pfem_amd.now_ns() - t_all) * 1e-9
    st = srv.stats()
    srv.close()
    n = len(stream)
    return dict(steps_per_s=n * passes / total, seconds=total, pass_seconds=pass_s,
                pass_steps_per_s=[n / t for t in pass_s], failed=int(fails), max_relres=relres,
                krylov_iters_pass1=int(kit1), ticks=st["ticks"], peak_inflight=st["peak_inflight"],
                arena_bytes=st["arena_bytes"], arena_peak=st["arena_peak"])


def reset(stream, cases, meshes, states):
    for req in stream:
        c = cases[req["i"]]
        c.set_state(states[req["i"]])
        c.set_positions(meshes[req["i"]].pos)


def accuracy(stream, plain, batched):
    """Largest per-field difference of the batched-mixed results from the plain-FP64 ones,
    relative to the field's largest magnitude (velocity, pressure, positions)."""
    worst = {"pspg_velocity": 0.0, "pspg_pressure": 0.0, "pspg_position": 0.0, "wc_velocity": 0.0,
             "wc_pressure": 0.0}
    for req in stream:
        i, d = req["i"], req["dim"]
        a, b = plain[i].get_state(), batched[i].get_state()
        N = plain[i].N
        rel = lambda x, y: float(np.abs(x - y).max() / max(np.abs(x).max(), 1e-300))  # noqa: E731
        k = "pspg" if req["kind"] == "PSPG" else "wc"
        worst[k + "_velocity"] = max(worst[k + "_velocity"], rel(a[:d * N], b[:d * N]))
        worst[k + "_pressure"] = max(worst[k + "_pressure"], rel(a[d * N:(d + 1) * N], b[d * N:(d + 1) * N]))
        if k == "pspg":
            worst["pspg_position"] = max(worst["pspg_position"],
                                         rel(plain[i].get_positions(), batched[i].get_positions()))
    return worst


# ------------------------------------------------------------------------------------------
# CPU baseline: the oracle on the host cores
# ------------------------------------------------------------------------------------------
def cpu_threads():
    """Host threads of the CPU baseline: OMP_NUM_THREADS when the environment sets it (the GPU
    box gives each GPU 16), else every core this process may run on."""
    env = os.environ.get("OMP_NUM_THREADS", "")
    if env.isdigit() and int(env) > 0:
        return int(env)
    return len(os.sched_getaffinity(0))


REF_TIMERS = {  # the reference's accumulated-time names (MomContEquationPSPG.inl:19-145,262-372)
    "t_triplets": "Compute triplets", "t_push": "Push back (n, n, 1)", "t_set_from_triplets": "Assemble matrix",
    "t_vector": "Assemble vector", "t_bc": "Apply boundary conditions",
    "t_solve": "Analyse pattern/Factorize/Solve system (Krylov stand-in)", "t_update": "Update solutions",
    "t_residual": "Compute Picard Algo residual"}


def _stats(ms):
    return dict(median_ms=pct(ms, 50), p99_ms=pct(ms, 99), mean_ms=float(np.mean(ms)), steps=len(ms))


def cpu_stream_sample(stream, meshes, states, threads, budget_s=20.0, seed=7):
    """Oracle (C++ FP64 restatement, BiCGSTAB + ILU(0) as the stand-in for SparseLU) on a
    seeded random sample of the stream's requests within about budget_s of CPU time, spread
    over the four kind/dimension groups; a request is taken only when its cost, estimated from
    the group's measured cost per node so far, fits the group's remaining budget.  The stream's
    CPU time is the per-node cost of each group times the group's nodes."""
    from oracle import oracle as O

    rng = np.random.default_rng(seed)
    groups = {}
    for req in stream:
        groups.setdefault((req["kind"], req["dim"]), []).append(req)
    per_node, sample_nodes, sample_s = {}, 0, 0.0
    sample = []
    t_all = time.perf_counter()
    for key, reqs in sorted(groups.items()):
        order = sorted(reqs, key=lambda r: meshes[r["i"]].n_nodes)
        first, rest = order[0], [order[k] for k in rng.permutation(np.arange(1, len(order)))]
        budget = budget_s / len(groups)
        spent, nodes = 0.0, 0
        for req in [first] + rest:
            m = meshes[req["i"]]
            if nodes > 0 and spent + spent / nodes * m.n_nodes > budget:
                continue
            q = states[req["i"]].copy()
            pos = m.pos.reshape(-1).copy()
            t0 = time.perf_counter()
            if req["kind"] == "PSPG":
                O.pspg_step(m, pos, q, dt=req["dt"], residual="Ax_f", min_res=1e-6, max_iter=10, solver_kind=2,
                            rtol=1e-8, nthreads=threads)
            else:
                d, N = m.dim, m.n_nodes
                v, p, rho, a = (q[:d * N].copy(), q[d * N:(d + 1) * N].copy(), q[(d + 1) * N:(d + 2) * N].copy(),
                                q[(d + 2) * N:].copy())
                O.wc_step(m, pos, v, p, rho, a, req["dt"], mu=1e-3, K0=K0, K0p=K0P, rho_star=RHO_STAR,
                          nthreads=threads)
                O.wc_next_dt(m, pos, v, p, rho, mu=1e-3, K0=K0, K0p=K0P, security_coeff=WC_SC, max_dt=WC_MAXDT,
                             nthreads=threads)
            spent += time.perf_counter() - t0
            nodes += m.n_nodes
            sample.append(req["i"])
        per_node[key] = spent / nodes
        sample_nodes += nodes
        sample_s += spent
    est = sum(per_node[(r["kind"], r["dim"])] * meshes[r["i"]].n_nodes for r in stream)
    sizes = sorted(meshes[i].n_nodes for i in sample)
    return {
        "value": len(stream) / est,
        "unit": "steps/s",
        "cores": threads,
        "kind": "port",
        "sample": (f"oracle C++ FP64 restatement (triplets, setFromTriplets-equivalent sum, Dirichlet elimination, "
                   f"Picard loop with BiCGSTAB+ILU(0) to relres 1e-8 as the SparseLU stand-in; WC CDS_dpdt + CFL) "
                   f"on {len(sample)} requests of the stream drawn at random (seed {seed}) within "
                   f"{budget_s:.0f} s ({sizes[0]}-{sizes[-1]} nodes, {sample_s:.1f} s measured, "
                   f"{len(sample) / sample_s:.2f} requests/s on the sample itself), scaled to the "
                   f"{len(stream)}-request stream by each kind/dimension's measured cost per node; "
                   f"OMP_PROC_BIND={os.environ.get('OMP_PROC_BIND')} OMP_PLACES={os.environ.get('OMP_PLACES')}"),
        "seconds_per_node": {f"{k}{d}D": v for (k, d), v in per_node.items()},
        "sample_requests": len(sample),
        "sample_requests_per_s": len(sample) / sample_s,
        "wall_s": time.perf_counter() - t_all,
    }


def cpu_configs(threads, wc_steps=3):
    """Measured CPU figures of the oracle on the fixed configurations, with the reference's
    phase names: the 2D dam break (1 warm-up + 10 Picard steps), one assembly + BC of the 3D
    dam break (262,701 nodes; its Krylov stand-in solve would take minutes per Picard iteration
    and is not run), and the 3D WC sloshing tank (2,060,451 nodes; 1 warm-up + wc_steps steps
    with computeNextDT)."""
    from oracle import oracle as O
    from pfem_amd import meshgen as mg

    out = {}
    # C1
    mesh = mg.dambreak_2d(h=0.05)
    q = mg.pspg_hydrostatic(mesh)
    pos = mesh.pos.reshape(-1).copy()
    O.pspg_step(mesh, pos.copy(), q.copy(), dt=1e-3, residual="Ax_f", min_res=1e-6, max_iter=10, solver_kind=2,
                rtol=1e-12, nthreads=threads)
    ms, phases = [], {v: 0.0 for v in REF_TIMERS.values()}
    for k in range(10):
        t0 = time.perf_counter()
        st, info = O.pspg_step(mesh, pos, q, dt=1e-3, residual="Ax_f", min_res=1e-6, max_iter=10, solver_kind=2,
                               rtol=1e-12, nthreads=threads)
        ms.append((time.perf_counter() - t0) * 1e3)
        for key, name in REF_TIMERS.items():
            phases[name] += info[key] * 1e3
    out["dambreak2d"] = dict(workload="2D PSPG dam break, 1,001 nodes, dt 1e-3, 10 steps (BiCGSTAB+ILU(0) to 1e-12)",
                             steps_per_s=1e3 * len(ms) / sum(ms), **_stats(ms),
                             phase_ms_total={k: v for k, v in phases.items()})
    # C2: one assembly + BC
    mesh = mg.dambreak_3d(h=0.01, col=(0.5, 0.5, 1.0), seed=0)
    q = mg.pspg_hydrostatic(mesh)
    t0 = time.perf_counter()
    S = O.pspg_assemble(mesh, q, q, dt=5e-4, nthreads=threads)
    wall = (time.perf_counter() - t0) * 1e3
    names = ["Compute triplets", "Push back (n, n, 1)", "Assemble matrix", "Assemble vector",
             "Apply boundary conditions"]
    out["dambreak3d_assembly"] = dict(workload=f"3D PSPG dam break, {mesh.n_nodes} nodes, {mesh.n_elems} tets: one "
                                               f"m_buildAbPSPG + m_applyBCPSPG",
                                      wall_ms=wall, elements_per_s=mesh.n_elems / (wall * 1e-3),
                                      phase_ms=dict(zip(names, (S.times * 1e3).tolist())), nnz=int(S.row_ptr[-1]))
    del S
    # C3
    mesh = mg.tank_3d(h=0.01)
    v, p, rho, a = mg.wc_tait_hydrostatic(mesh, K0=K0, K0p=K0P, rho_star=RHO_STAR)
    v, a = v.reshape(-1).copy(), a.reshape(-1).copy()
    pos = mesh.pos.reshape(-1).copy()
    dt = O.wc_next_dt(mesh, pos, v, p, rho, mu=1e-3, K0=K0, K0p=K0P, security_coeff=WC_SC, max_dt=WC_MAXDT,
                      nthreads=threads)
    ms, tc, tm, tdt = [], 0.0, 0.0, 0.0
    for k in range(wc_steps + 1):
        t0 = time.perf_counter()
        tt = O.wc_step(mesh, pos, v, p, rho, a, dt, mu=1e-3, K0=K0, K0p=K0P, rho_star=RHO_STAR, nthreads=threads)
        t1 = time.perf_counter()
        dt = O.wc_next_dt(mesh, pos, v, p, rho, mu=1e-3, K0=K0, K0p=K0P, security_coeff=WC_SC, max_dt=WC_MAXDT,
                          nthreads=threads)
        t2 = time.perf_counter()
        if k == 0:
            continue
        ms.append((t2 - t0) * 1e3)
        tc += tt[0] * 1e3
        tm += tt[1] * 1e3
        tdt += (t2 - t1) * 1e3
    out["wc3d"] = dict(workload=f"3D WC sloshing tank, {mesh.n_nodes} nodes, {mesh.n_elems} tets, CDS_dpdt + "
                                f"computeNextDT, {wc_steps} steps after 1 warm-up",
                       steps_per_s=1e3 * len(ms) / sum(ms), **_stats(ms),
                       phase_ms_total={"Solving continuity eq": tc, "Solving momentum eq": tm,
                                       "computeNextDT": tdt})
    return out


# ------------------------------------------------------------------------------------------
# the 2D dam break that the reference itself runs on the CPU
# ------------------------------------------------------------------------------------------
def run_dambreak2d(ctx, steps=10, threads=1, with_cpu=True):
    """2D PSPG dam break (1.0 x 2.0 m column in a 4.0 x 3.0 m tank, h = 0.05 m, 1,001 nodes):
    `steps` Picard steps of dt 1e-3 through the plain FP64 drop-in, and the same steps through
    the oracle on the host (BiCGSTAB + ILU(0) to relres 1e-12 standing in for SparseLU; its
    dense LU takes ~9 s a step here); the largest state and position differences after the
    last step."""
    import pfem_amd
    from pfem_amd import meshgen as mg

    dt = 1e-3
    mesh = mg.dambreak_2d(h=0.05)
    q0 = mg.pspg_hydrostatic(mesh)

    def steps_of(graph):
        """`steps` timed steps from q0 on a fresh case; graph: BiCGSTAB chunks replayed from
        captured HIP graphs (PFEM_GRAPH=1) instead of launched kernel by kernel."""
        old = os.environ.get("PFEM_GRAPH")
        os.environ["PFEM_GRAPH"] = "1" if graph else "0"
        eq = pfem_amd.MomContEqIncompNewton(ctx, 2, rho=1000.0, mu=1e-3, residual="Ax_f", min_res=1e-6,
                                            max_iter=10, precision="fp64", krylov_rtol=1e-12, krylov_max_iter=20000)
        eq.set_mesh(mesh)
        eq.set_state(q0)
        eq.solve(dt)  # warm-up step (kernel loading), then restart from the initial state
        eq.set_state(q0)
        eq.set_positions(mesh.pos)
        lat = []
        t0 = time.perf_counter()
        for k in range(steps):
            t1 = time.perf_counter()
            if not eq.solve(dt):
                raise SystemExit(f"2D dam break step {k} failed: {eq.last_stats}")
            lat.append((time.perf_counter() - t1) * 1e3)
        total = time.perf_counter() - t0
        res = (eq.get_state(), eq.get_positions().reshape(-1), total, lat, eq.engine_stats())
        eq.close()
        if old is None:
            del os.environ["PFEM_GRAPH"]
        else:
            os.environ["PFEM_GRAPH"] = old
        return res

    q_gpu, pos_gpu, total, lat, _ = steps_of(False)
    qg_graph, pg_graph, total_g, lat_g, gst = steps_of(True)
    out = dict(workload=f"2D PSPG dam break, plain FP64 drop-in ({mesh.n_nodes} nodes, {mesh.n_elems} triangles, "
                        f"dt {dt}, {steps} steps)",
               steps_per_s=steps / total, p50_ms=pct(lat, 50), p99_ms=pct(lat, 99),
               graph=dict(steps_per_s=steps / total_g, p50_ms=pct(lat_g, 50), p99_ms=pct(lat_g, 99),
                          same_bits=bool(np.array_equal(qg_graph, q_gpu) and np.array_equal(pg_graph, pos_gpu)),
                          **gst))
    if with_cpu:
        from oracle import oracle as O

        pos = mesh.pos.reshape(-1).copy()
        q = q0.copy()
        t0 = time.perf_counter()
        for k in range(steps):
            st, _ = O.pspg_step(mesh, pos, q, dt=dt, residual="Ax_f", min_res=1e-6, max_iter=10, solver_kind=2,
                                rtol=1e-12, nthreads=threads)
            if st != 0:
                raise SystemExit(f"oracle 2D dam break step {k} failed")
        cpu = time.perf_counter() - t0
        out.update(oracle_steps_per_s=steps / cpu, oracle_threads=threads,
                   max_state_diff_vs_oracle=float(np.abs(q_gpu - q).max() / np.abs(q).max()),
                   max_position_diff_vs_oracle=float(np.abs(pos_gpu - pos).max()))
    return out


# ------------------------------------------------------------------------------------------
# the 3D weakly compressible sloshing tank and the mixed-precision 3D column field
# ------------------------------------------------------------------------------------------
def wc_step_bytes(dim, nodes, elems, s_state):
    """Algorithmic bytes of one WC step with its computeNextDT (folded into the step's kernels):
    connectivity read by the two element passes (continuity, momentum), the node state v, a
    (dim each), p, rho read and written once, positions read and written once, and the FP64
    accumulators (continuity mass and rhs, each node's smallest element size for the CFL,
    momentum mass, dim force components) written and read once."""
    conn = 2 * elems * (dim + 1) * 4
    state = 2 * nodes * (2 * dim + 2) * s_state
    pos = 2 * nodes * dim * 8
    acc = 2 * nodes * (4 + dim) * 8
    return conn + state + pos + acc


def run_wc3d(ctx, steps, warmup=3):
    """3D weakly compressible sloshing: CDS_dpdt steps with computeNextDT, FP32 node state and
    FP64 accumulators, on the 2 x 2 x 1 m tank half filled at h = 0.01 m."""
    import pfem_amd
    from pfem_amd import meshgen as mg

    mesh = mg.tank_3d(h=0.01)
    v, p, rho, a = mg.wc_tait_hydrostatic(mesh, K0=K0, K0p=K0P, rho_star=RHO_STAR)
    out = {}
    for state in ("fp32", "fp64"):
        case = pfem_amd.WCompNewtonStep(ctx, 3, mu=1e-3, K0=K0, K0p=K0P, rho_star=RHO_STAR, wc_state=state)
        case.set_mesh(mesh)
        case.set_state(np.concatenate([v.reshape(-1), p, rho, a.reshape(-1)]))
        dt = case.computeNextDT(WC_SC, WC_MAXDT)
        for _ in range(warmup):
            case.solveOneTimeStep(dt)
            dt = case.computeNextDT(WC_SC, WC_MAXDT)
        lat = []
        t0 = time.perf_counter()
        for _ in range(steps):
            t1 = time.perf_counter()
            case.solveOneTimeStep(dt)
            dt = case.computeNextDT(WC_SC, WC_MAXDT)
            lat.append((time.perf_counter() - t1) * 1e3)
        total = time.perf_counter() - t0
        sb = 4 if state == "fp32" else 8
        b = wc_step_bytes(3, mesh.n_nodes, mesh.n_elems, sb)
        # the extractors (MassExtractor, MinMaxExtractor) on the stepped tank, outside the clock
        # of the steps: host wall time per call, including the readback of the partials
        mass = case.mass()
        t2 = time.perf_counter()
        for _ in range(10):
            case.mass()
        mass_ms = (time.perf_counter() - t2) * 100.0
        t3 = time.perf_counter()
        for _ in range(10):
            zmax = case.minmax(2, "max")
        minmax_ms = (time.perf_counter() - t3) * 100.0
        out[state] = dict(steps_per_s=steps / total, p50_ms=pct(lat, 50), p99_ms=pct(lat, 99),
                          elements_per_s=mesh.n_elems * steps / total, dt_last=dt,
                          algorithmic_bytes_per_step=b, achieved_GBps=b * steps / total / 1e9,
                          frac_hbm_peak=b * steps / total / 1e9 / HBM_PEAK_GBPS,
                          extract=dict(mass=mass, mass_ms=mass_ms, z_max=zmax, minmax_ms=minmax_ms))
        case.close()
    return dict(workload=f"3D WC sloshing tank 2x2x1 m half filled, h=0.01 m ({mesh.n_nodes} nodes, "
                         f"{mesh.n_elems} tets), CDS_dpdt + computeNextDT (securityCoeff {WC_SC})",
                steps=steps, **out)


def run_precision3d(ctx, steps):
    """3D PSPG column field (seeded random-walk heights, ~1.07M nodes): FP32 inner BiCGSTAB +
    block Jacobi (inner relres 3e-2, at most 200 iterations per sweep) inside FP64 iterative
    refinement to relres 1e-10, against the FP64 BiCGSTAB
    solve to the same relres; per-step largest error of u, p and the free-surface positions,
    relative to each field's largest magnitude."""
    import pfem_amd
    from pfem_amd import meshgen as mg

    mesh = mg.randomwalk_3d(h=0.01, ncol=100, mean_cells=100, seed=0)
    q0 = mg.pspg_hydrostatic(mesh)
    N = mesh.n_nodes
    dt = 5e-4
    runs = {}
    for prec in ("fp64", "mixed_ir"):
        # inner FP32 solves stop at 3e-2 or 200 iterations (measured over these steps: total
        # 4.6 s, against 5.9 s at 1e-3 / 400 and 7.0 s for the FP64 solve)
        c = pfem_amd.MomContEqIncompNewton(ctx, 3, rho=1000.0, mu=1e-3, residual="Ax_f", min_res=1e-6, max_iter=10,
                                           precision=prec, precond="jacobi", krylov_rtol=1e-10,
                                           krylov_max_iter=40000, inner_rtol=3e-2, inner_max_iter=200)
        c.set_mesh(mesh)
        c.set_state(q0)
        qs, ps, lat, st = [], [], [], []
        for k in range(steps):
            t1 = time.perf_counter()
            ok = c.solve(dt)
            lat.append((time.perf_counter() - t1) * 1e3)
            if not ok:
                raise SystemExit(f"precision case {prec} step {k} failed: {c.last_stats}")
            st.append(c.last_stats)
            qs.append(c.get_state())
            ps.append(c.get_positions())
        runs[prec] = dict(q=qs, pos=ps, lat=lat, st=st)
        c.close()
    surf = np.zeros(N, bool)
    top = mesh.surface - mesh.pos[2] < 0.5 * mesh.h
    surf[top] = True
    err = []
    for k in range(steps):
        q64, q32 = runs["fp64"]["q"][k], runs["mixed_ir"]["q"][k]
        u64, u32 = q64[:3 * N], q32[:3 * N]
        p64, p32 = q64[3 * N:], q32[3 * N:]
        x64 = runs["fp64"]["pos"][k].reshape(3, N)[:, surf]
        x32 = runs["mixed_ir"]["pos"][k].reshape(3, N)[:, surf]
        err.append(dict(velocity=float(np.abs(u32 - u64).max() / max(np.abs(u64).max(), 1e-300)),
                        pressure=float(np.abs(p32 - p64).max() / max(np.abs(p64).max(), 1e-300)),
                        free_surface_position=float(np.abs(x32 - x64).max() / mesh.h)))
    summ = lambda r: dict(p50_ms=pct(r["lat"], 50), mean_ms=float(np.mean(r["lat"])),
                          total_s=float(np.sum(r["lat"]) * 1e-3),
                          krylov_iters_per_step=float(np.mean([s["krylov_iters"] for s in r["st"]])),
                          max_relres=float(max(s["max_relres"] for s in r["st"])))
    return dict(workload=f"3D PSPG random-walk column field, {N} nodes, {mesh.n_elems} tets, dt {dt}, "
                         f"linear solves to relres 1e-10",
                fp64=summ(runs["fp64"]), mixed_ir=summ(runs["mixed_ir"]),
                mixed_ir_speedup=float(np.sum(runs["fp64"]["lat"]) / np.sum(runs["mixed_ir"]["lat"])),
                error_vs_fp64_per_step=err,
                error_note="velocity and pressure: max abs difference / field max; free-surface position: "
                           "max abs difference of the top-layer nodes in units of h")


# ------------------------------------------------------------------------------------------
# the 3D dam break drop-in and the SpMV roofline
# ------------------------------------------------------------------------------------------
def spmv_fused_bytes(nodes, n, nnz, graph_nnz, vbytes=8):
    """Algorithmic bytes of the first BiCGSTAB SpMV (v = A ph; dot (rh, v)) in the node-block
    format: the nnz values once, the node graph once (neighbour indices, pointers and the
    16-byte per-node row descriptor), ph read once, v written once, rh read once."""
    return nnz * vbytes + graph_nnz * 4 + (nodes + 1) * 4 + nodes * 16 + 3 * n * vbytes


def pmc_traffic(path=os.path.join(ROOT, "profiles", "spmv_kry_v_dambreak3d_262k_pmc_traffic.md")):
    """HBM bytes per launch of the dam-break SpMV from the committed rocprofv3 FETCH_SIZE /
    WRITE_SIZE passes (scaled by the counters' readings for a copy kernel of known size);
    None when the summary is absent.  The counters need the profiler, so they cannot be read
    inside this process."""
    try:
        with open(path) as f:
            for line in f:
                if line.startswith("traffic_bytes_per_launch:"):
                    return int(line.split(":")[1])
    except OSError:
        pass
    return None


def graph_nnz_from_csr(rp, dim, nodes):
    """Node-graph entries: a node's continuity row holds (dim+1)*deg entries (1 for a free node)."""
    cont = np.diff(rp[dim * nodes:])
    return int(np.sum(cont // (dim + 1)))


def run_dambreak3d(ctx, steps, warmup):
    import pfem_amd
    from pfem_amd import meshgen as mg

    dt = 5e-4
    mesh = mg.dambreak_3d(h=0.01, col=(0.5, 0.5, 1.0), seed=0)
    q0 = mg.pspg_hydrostatic(mesh)
    case = pfem_amd.MomContEqIncompNewton(ctx, 3, rho=1000.0, mu=1e-3, residual="Ax_f", min_res=1e-6, max_iter=10,
                                          precision="fp64", krylov_rtol=1e-9, krylov_max_iter=20000)
    case.set_mesh(mesh)
    case.set_state(q0)
    for w in range(max(1, warmup)):
        ok = case.solve(dt)
        log(f"[bench] dam break 3D warmup {w}: ok={ok} {case.last_stats}")
    rp, _, _ = case.csr()
    n, nnz = len(rp) - 1, int(rp[-1])
    case.profile_spmv(True)
    lat, stats = [], []
    t0 = time.perf_counter()
    for k in range(steps):
        t1 = time.perf_counter()
        ok = case.solve(dt)
        lat.append((time.perf_counter() - t1) * 1e3)
        stats.append(case.last_stats)
        if not ok:
            raise SystemExit(f"dam break step {k} failed: {case.last_stats}")
    total = time.perf_counter() - t0
    prof = case.profile_spmv(False)
    # the same SpMV kernel replayed back to back on the last step's system (its duration
    # without the event packets that the per-dispatch timing inside the solve adds), and
    # the ceilings on the same device memory: a streaming read of the values and a copy
    # kernel (also the known-size kernel the HBM byte counters are scaled by)
    # (five groups of 50 launches; the median group, as one group's average varies by ~2 %)
    groups = [case.bench_kernels(50) for _ in range(5)]
    kb = dict(groups[0])
    for key in ("kry_v_us", "stream_read_us", "copy_us", "amg_res0_us"):
        kb[key] = float(np.median([g[key] for g in groups]))
    gnnz = graph_nnz_from_csr(rp, 3, mesh.n_nodes)
    bytes_spmv = spmv_fused_bytes(mesh.n_nodes, n, nnz, gnnz)
    spmv_ms = kb["kry_v_us"] * 1e-3
    achieved = bytes_spmv / (spmv_ms * 1e-3) / 1e9 if spmv_ms > 0 else 0.0
    t_asm = float(np.mean([s["t_assemble"] for s in stats]))
    out = dict(
        workload="3D PSPG dam break, plain FP64 drop-in (column 0.5x0.5x1.0 m, h=0.01 m, seeded jitter)",
        nodes=mesh.n_nodes, tets=mesh.n_elems, unknowns=n, nnz=nnz, graph_nnz=gnnz, dt=dt,
        steps_per_s=steps / total, p50_ms=pct(lat, 50), p99_ms=pct(lat, 99),
        picard_iters_per_step=float(np.mean([s["picard_iters"] for s in stats])),
        krylov_iters_per_step=float(np.mean([s["krylov_iters"] for s in stats])),
        assemblies_per_step=float(np.mean([s["assemblies"] for s in stats])),
        max_relres=float(max(s["max_relres"] for s in stats)),
        elements_per_s=mesh.n_elems / t_asm if t_asm > 0 else None,
    )
    roof = {
        "kernel": "k_kry_v<double,16> (BiCGSTAB SpMV v = A ph over the sliced ELL copy, fused (rh, v)), "
                  "3D dam break 262,701 nodes",
        "bound": "hbm",
        "achieved": achieved,
        "peak": HBM_PEAK_GBPS,
        "unit": "GB/s",
        "frac": achieved / HBM_PEAK_GBPS,
        "traffic": pmc_traffic(),
        "traffic_source": "profiles/spmv_kry_v_dambreak3d_262k_pmc_traffic.md (rocprofv3 FETCH_SIZE x 1.9988 + "
                          "WRITE_SIZE x 1.0 over the replay launches, factors from a copy kernel and a "
                          "non-temporal streaming read of known size)",
        "avg_launch_ms": spmv_ms,
        "launch_timing": "HIP events around 50 back-to-back launches of the kernel on the case's stream, "
                         "after the timed steps; median of 5 such groups",
        "launch_groups_ms": [g["kry_v_us"] * 1e-3 for g in groups],
        "in_solve_event_ms": prof["avg_ms_fp64"],
        "in_solve_launches_timed": prof["launches_fp64"],
        "algorithmic_bytes_per_launch": bytes_spmv,
        "ceilings_GBps": {"stream_read": nnz * 8 / kb["stream_read_us"] / 1e3,
                          "copy": kb["copy_bytes"] / kb["copy_us"] / 1e3},
    }
    if kb["amg_res0_us"] > 0:
        # the kernel that takes the most time in this case's steps: the multigrid's level-0
        # residual t = r - A w (the same ELL values, node graph and three vectors as the SpMV)
        gbps = bytes_spmv / (kb["amg_res0_us"] * 1e-6) / 1e9
        roof["busiest_kernel"] = {
            "kernel": "k_amg_res<double> at level 0 (multigrid residual t = r - A w over the FP64 ELL copy)",
            "achieved": gbps, "peak": HBM_PEAK_GBPS, "unit": "GB/s", "frac": gbps / HBM_PEAK_GBPS,
            "avg_launch_ms": kb["amg_res0_us"] * 1e-3, "algorithmic_bytes_per_launch": bytes_spmv,
            "launch_timing": "HIP events around 50 back-to-back launches as a V-cycle launches it; median of 5 groups",
        }
    out["remesh"] = run_remesh_cycle(case, mesh, q0, dt)
    case.close()
    return out, roof


def run_remesh_cycle(case, mesh, q0, dt):
    """What the reference does after every successful step (IncompNewton/Solver.cpp:236-246,
    Mesh.cpp:919-926), on the C2 nodes: the host mesher's alpha-shape triangulation of the node
    cloud, the new mesh's set-up on the GPU (pattern, ELL, multigrid), a step on it, then one
    full remesh (pfem_case_remesh: nodes to the host, smoothing / add / remove / Delaunay /
    alpha shape, set-up) and the next step.  Mesh parameters of the 3D Koshizuka dam break
    (examples/3D/damBreakKoshizuka/damBreakKoshizuka3DComp.lua)."""
    import pfem_amd

    m = pfem_amd.Mesher(3, hchar=mesh.h, alpha=1.2, omega=0.7, gamma=0.35, addOnFS=True, deleteFlyingNodes=False)
    m.set_nodes(mesh.pos, mesh.flags, q0.reshape(4, -1))
    t0 = time.perf_counter()
    m.triangulate()
    t_tri = time.perf_counter() - t0
    st_tri = m.stats()
    t0 = time.perf_counter()
    case.set_mesh_from(m)
    t_set0 = time.perf_counter() - t0
    setup0 = case.setup_times()
    t0 = time.perf_counter()
    ok1 = case.solve(dt)
    t_step1 = time.perf_counter() - t0
    t0 = time.perf_counter()
    case.remesh(m)
    t_rem = time.perf_counter() - t0
    st = m.stats()
    setup1 = case.setup_times()
    t0 = time.perf_counter()
    ok2 = case.solve(dt)
    t_step2 = time.perf_counter() - t0
    n, e, _ = m.sizes()
    log(f"[bench] remesh: triangulate {t_tri:.3f}s set-up {t_set0:.3f}s step {t_step1 * 1e3:.1f}ms "
        f"remesh {t_rem:.3f}s ({st}) step {t_step2 * 1e3:.1f}ms ok={ok1},{ok2}")
    return dict(
        triangulate_s=t_tri, triangulate_delaunay_s=st_tri["t_delaunay"], triangulate_alpha_s=st_tri["t_alpha"],
        set_mesh_s=t_set0, set_mesh_phases_s=setup0, first_step_ms=t_step1 * 1e3, first_step_ok=bool(ok1),
        remesh_call_s=t_rem, remesh_host_s={"nodes": st["t_nodes"], "delaunay": st["t_delaunay"],
                                            "alpha_shape": st["t_alpha"]},
        remesh_set_mesh_phases_s=setup1, nodes_added=st["added"], nodes_removed=st["removed"],
        nodes_after=n, elements_after=e, next_step_ms=t_step2 * 1e3, next_step_ok=bool(ok2),
        note="host mesher single-threaded; pfem_case_remesh = positions/states download + host remesh + set_mesh "
             "+ state upload",
    )


# ------------------------------------------------------------------------------------------
# ------------------------------------------------------------------------------------------
# the other solvers of the path: heat (Conduction), fractional step, Boussinesq
# ------------------------------------------------------------------------------------------
def run_secondary(ctx, steps=10, with_cpu=True, threads=1):
    """Steps of the heat equation (3D cavity, 40^3 cells), the fractional-step solver (3D
    lid-driven cavity, 24^3 cells) and the Boussinesq problem (2D heated cavity, 100^2 cells) on
    the GPU, each with the oracle's time for one step of the same case on the host."""
    import pfem_amd
    from pfem_amd import meshgen as mg

    out = {}

    def timed(fn, n):
        ms = []
        for _ in range(n):
            t0 = time.perf_counter()
            fn()
            ms.append((time.perf_counter() - t0) * 1e3)
        return ms

    # heat: bottom 360 K / top 300 K, insulated sides, steel-like material, dt 0.5 s
    m = mg.cavity(3, 40)
    N = m.n_nodes
    fl = np.where(m.meta["top"] | m.meta["bottom"], mg.DIRICHLET | mg.FIXED, mg.FIXED).astype(np.uint8)
    hm = mg.PfemMesh(3, m.pos, m.conn, fl, m.h)
    Tbc = np.where(m.meta["bottom"], 360.0, np.where(m.meta["top"], 300.0, 0.0))
    T0 = np.where(m.meta["bottom"], 360.0, np.where(m.meta["top"], 300.0, 330.0))
    mat = dict(rho=8000.0, cv=500.0, k=50.0)
    dt = 0.5
    hc = pfem_amd.HeatEqIncompNewton(ctx, 3, krylov_rtol=1e-10, **mat)
    hc.set_mesh(hm)
    hc.set_state(T0)
    hc.set_dirichlet_T(Tbc)
    hc.solve(dt)
    iters = []

    def heat_step():
        assert hc.solve(dt)
        iters.append(hc.last_stats["krylov_iters"])
    ms = timed(heat_step, steps)
    res = dict(workload=f"3D heat conduction, cavity 40^3 cells ({N} nodes, {m.n_elems} tets), Dirichlet bottom / "
                        f"top, rho 8000 cv 500 k 50, dt {dt} s, CG to relres 1e-10",
               steps_per_s=1e3 * len(ms) / sum(ms), **_stats(ms), cg_iters_per_step=float(np.mean(iters)))
    if with_cpu:
        from oracle import oracle as O
        T = T0.copy()
        t0 = time.perf_counter()
        O.heat_step(hm, T, Tbc=Tbc, hc=O.HeatCase(**mat), dt=dt, rtol=1e-10)
        res["oracle_ms_per_step"] = (time.perf_counter() - t0) * 1e3
        res["oracle_threads"] = threads
    hc.close()
    out["heat3d"] = res
    log(f"[bench] heat 3D: {res}")

    # fractional step: lid (1, 0, 0) on the top wall, gammaFS 1, residual U
    m = mg.cavity(3, 24)
    N = m.n_nodes
    vbc = np.zeros((3, N))
    vbc[0, m.meta["lid"]] = 1.0
    q0 = np.zeros(4 * N)
    q0[:3 * N] = vbc.reshape(-1)
    kw = dict(rho=1000.0, mu=1e-2, body_force=[0.0, 0.0, -9.81], gammaFS=1.0, residual="U", min_res=1e-6,
              max_iter=10)
    dt = 1e-3
    fc = pfem_amd.MomContEqFracStep(ctx, 3, krylov_rtol=1e-10, **kw)
    fc.set_mesh(m)
    fc.set_state(q0)
    fc.set_dirichlet(vbc)
    stats = []

    def frac_step():
        rc, st = fc.step(dt)
        assert rc == 0, st
        stats.append(st)
    frac_step()
    stats.clear()
    ms = timed(frac_step, steps)
    res = dict(workload=f"3D lid-driven cavity, fractional step, 24^3 cells ({N} nodes, {m.n_elems} tets), "
                        f"rho 1000 mu 1e-2, dt {dt}, Picard U 1e-6, CG to relres 1e-10",
               steps_per_s=1e3 * len(ms) / sum(ms), **_stats(ms),
               picard_iters_per_step=float(np.mean([x["picard_iters"] for x in stats])),
               cg_iters_per_step=float(np.mean([x["krylov_iters"] for x in stats])))
    if with_cpu:
        from oracle import oracle as O
        q, pos = q0.copy(), m.pos.copy()
        t0 = time.perf_counter()
        O.frac_step(m, pos, q, dir_vals=vbc.reshape(-1).copy(), rho=1000.0, mu=1e-2, g=[0.0, 0.0, -9.81], dt=dt,
                    gamma_fs=1.0, residual="U", min_res=1e-6, max_iter=10, rtol=1e-10)
        res["oracle_ms_per_step"] = (time.perf_counter() - t0) * 1e3
        res["oracle_threads"] = threads
    fc.close()
    out["frac3d"] = res
    log(f"[bench] FracStep 3D: {res}")

    # Boussinesq: a fluid layer heated from below (360 K), walls at rest, an open top (free
    # surface; a closed cavity leaves the PSPG pressure without a level, a singular system)
    m = mg.cavity(2, 100)
    N = m.n_nodes
    fl = m.flags.copy()
    fl[m.meta["lid"]] = mg.ONFS
    fl[m.meta["bottom"]] |= 32  # PFEM_NODE_DIRICHLET_T
    bm = mg.PfemMesh(2, m.pos, m.conn, fl, m.h)
    Tbc = np.where(m.meta["bottom"], 360.0, 0.0)
    rng = np.random.default_rng(3)
    T0 = np.where(m.meta["bottom"], 360.0, 330.0 + rng.normal(scale=2.0, size=N))
    g = [0.0, -9.81]
    q0 = np.concatenate([np.zeros(2 * N), 1000.0 * 9.81 * (1.0 - m.pos[1]), T0])
    bmat = dict(rho=1000.0, mu=1e-2, cv=4000.0, k=50.0, alpha=2e-3, Tr=330.0)
    dt = 5e-3
    bc = pfem_amd.BoussinesqIncompNewton(ctx, 2, body_force=g, residual="U", min_res=1e-6, max_iter=10,
                                         krylov_rtol=1e-10, **bmat)
    bc.set_mesh(bm)
    bc.set_state(q0)
    bc.set_dirichlet(np.zeros((2, N)))
    bc.set_dirichlet_T(Tbc)

    def bq_step():
        assert bc.solve(dt), bc.failed
    bq_step()
    ms = timed(bq_step, steps)
    res = dict(workload=f"2D fluid layer heated from below, open top, Boussinesq (PSPG + heat, heat first), "
                        f"100^2 cells ({N} nodes, {m.n_elems} triangles), alpha 2e-3, dt {dt}, Picard U 1e-6, "
                        f"solves to relres 1e-10",
               steps_per_s=1e3 * len(ms) / sum(ms), **_stats(ms))
    if with_cpu:
        from oracle import oracle as O
        q, pos = q0.copy(), m.pos.copy()
        t0 = time.perf_counter()
        O.boussinesq_step(bm, pos, q, O.HeatCase(rho=1000.0, cv=4000.0, k=50.0), 2e-3, 330.0, Tbc=Tbc, rho=1000.0,
                          mu=1e-2, g=g, dt=dt, residual="U", min_res=1e-6, max_iter=10, rtol=1e-10)
        res["oracle_ms_per_step"] = (time.perf_counter() - t0) * 1e3
        res["oracle_threads"] = threads
    bc.close()
    out["boussinesq2d"] = res
    log(f"[bench] Boussinesq 2D: {res}")
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5,
                    help="timed passes of the serving stream (each pass: every request once, one time step each)")
    ap.add_argument("--warmup", type=int, default=1, help="untimed warm-up passes of the batched path")
    ap.add_argument("--stream-cases", type=int, default=512, help="requests in the serving stream")
    # slots: 64 / 128 / 192 / 256 / 384 served the 512-request stream at 144 / 156 / 161 / 166 /
    # 166 steps/s (larger launches, shorter tail); the arena peaks near 59 GiB from 256 on
    ap.add_argument("--slots", type=int, default=256, help="server requests in flight at once")
    ap.add_argument("--budget-gb", type=float, default=120.0, help="server workspace arena (GiB)")
    ap.add_argument("--dambreak-steps", type=int, default=20)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--batched-precision", choices=["fp64", "mixed", "mixed_ir"], default="mixed",
                    help="PSPG solve precision of the batched path")
    ap.add_argument("--wc-steps", type=int, default=200, help="steps of the 3D WC sloshing tank")
    ap.add_argument("--precision-steps", type=int, default=3, help="steps of the 3D mixed-precision column field")
    ap.add_argument("--secondary-steps", type=int, default=10,
                    help="steps of the heat, fractional-step and Boussinesq cases")
    ap.add_argument("--only", choices=["batched", "plain", "dambreak", "dambreak2d", "wc3d", "precision",
                                       "secondary"],
                    default=None,
                    help="run one workload once and print its result only (for profilers)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: --gpus {args.gpus} is not supported (cases run whole on one GPU; "
              "spreading cases or partitions over several GPUs is a design only)", file=sys.stderr)
        sys.exit(2)
    import pfem_amd

    ctx = pfem_amd.Context(0)
    xcd = ctx.xcd_map(2048)
    xcd_round_robin = float(np.mean(xcd == (np.arange(xcd.size) % 8)))
    log(f"[bench] workgroup g on XCD g mod 8 for {xcd_round_robin:.3f} of 2048 workgroups")
    if args.only == "dambreak":
        db, roof = run_dambreak3d(ctx, args.dambreak_steps, 2)
        print(json.dumps({"dambreak3d_fp64": db, "roofline": roof}), flush=True)
        return
    if args.only == "dambreak2d":
        print(json.dumps({"dambreak2d_fp64": run_dambreak2d(ctx)}), flush=True)
        return
    if args.only == "wc3d":
        print(json.dumps({"wc3d": run_wc3d(ctx, args.wc_steps)}), flush=True)
        return
    if args.only == "precision":
        print(json.dumps({"precision3d": run_precision3d(ctx, args.precision_steps)}), flush=True)
        return
    if args.only == "secondary":
        print(json.dumps({"secondary": run_secondary(ctx, args.secondary_steps, with_cpu=not args.no_cpu_baseline,
                                                     threads=args.cpu_threads or cpu_threads())}), flush=True)
        return
    n_req = max(1, args.stream_cases)
    stream = build_stream(n_req)
    t0 = time.perf_counter()
    meshes = {r["i"]: make_mesh(r) for r in stream}
    states = {r["i"]: initial_state(r, meshes[r["i"]]) for r in stream}
    log(f"[bench] stream: {n_req} requests, {sum(m.n_nodes for m in meshes.values())} nodes, "
        f"{sum(m.n_elems for m in meshes.values())} elements; meshes in {time.perf_counter() - t0:.1f} s")
    t0 = time.perf_counter()
    plain = {r["i"]: make_case(ctx, r, meshes[r["i"]], "plain") for r in stream}
    batched = {r["i"]: make_case(ctx, r, meshes[r["i"]], "batched", args.batched_precision) for r in stream}
    for r in stream:
        plain[r["i"]].reserve_workspace()
        if r["kind"] == "PSPG":
            r["dt"] = 1e-3 if r["dim"] == 2 else 5e-4
        else:
            c = plain[r["i"]]
            c.set_state(states[r["i"]])
            r["dt"] = c.wc_next_dt(WC_SC, WC_MAXDT)
    dt0 = {r["i"]: r["dt"] for r in stream}
    log(f"[bench] {2 * n_req} cases resident in {time.perf_counter() - t0:.1f} s")
    budget = int(args.budget_gb * (1 << 30))
    passes, warm_passes = max(1, args.steps), max(0, args.warmup)

    def reset_all(cases):
        reset(stream, cases, meshes, states)
        for r in stream:
            r["dt"] = dt0[r["i"]]

    if args.only:
        cases = batched if args.only == "batched" else plain
        reset_all(cases)
        # the XCD probe's one-workgroup-per-slot kernel brackets the timed window in a kernel
        # trace, so a profile can be cut to the dispatches of the timed run (the cases' mesh
        # set-up comes before it)
        ctx.xcd_map(8)
        res = (run_batched(ctx, stream, cases, False, budget, args.slots) if args.only == "batched" else
               run_plain(stream, cases, False))
        ctx.xcd_map(8)
        print(json.dumps({"path": args.only, "requests": n_req, **res}), flush=True)
        return

    # plain FP64: a short warm-up (kernel loading), then one saturated pass of the stream from
    # the initial states
    reset_all(plain)
    run_plain(stream[:8], plain, poisson=False)
    reset_all(plain)
    plain_sat = run_plain(stream, plain, poisson=False)
    log(f"[bench] plain FP64 saturated pass: {plain_sat}")
    # batched mixed: warm-up passes (the first one from the same initial states gives the
    # accuracy against the plain pass), then `passes` timed passes from the initial states,
    # each advancing every case by one more time step
    reset_all(batched)
    acc = None
    for w in range(warm_passes):
        wres = run_batched(ctx, stream, batched, False, budget, args.slots)
        if w == 0:
            acc = accuracy(stream, plain, batched)
            log(f"[bench] batched mixed vs plain FP64 (same initial states, one step): {acc}")
        log(f"[bench] batched warm-up pass {w}: {wres['steps_per_s']:.1f} steps/s")
    reset_all(batched)
    timed = run_batched_passes(ctx, stream, batched, passes, budget, args.slots)
    log(f"[bench] batched mixed, {passes} timed passes: {timed['steps_per_s']:.1f} steps/s "
        f"(pass 1 {timed['pass_steps_per_s'][0]:.1f}), failed {timed['failed']}")
    if acc is None:
        acc = accuracy(stream, plain, batched) if passes == 1 else None
    # Poisson arrivals at RATE requests/s (latency from arrival to result), one pass each
    reset_all(plain)
    plain_poi = run_plain(stream, plain, poisson=True)
    log(f"[bench] plain FP64 poisson: {plain_poi}")
    reset_all(batched)
    batched_poi = run_batched(ctx, stream, batched, True, budget, args.slots)
    log(f"[bench] batched mixed poisson: {batched_poi}")
    for c in list(plain.values()) + list(batched.values()):
        c.close()

    db, roof = run_dambreak3d(ctx, args.dambreak_steps, 2)
    log(f"[bench] dam break 3D: {db}")
    db2 = run_dambreak2d(ctx, 10, with_cpu=not args.no_cpu_baseline, threads=cpu_threads())
    log(f"[bench] dam break 2D: {db2}")
    wc3d = run_wc3d(ctx, args.wc_steps) if args.wc_steps > 0 else None
    log(f"[bench] WC sloshing 3D: {wc3d}")
    prec3d = run_precision3d(ctx, args.precision_steps) if args.precision_steps > 0 else None
    log(f"[bench] precision 3D: {prec3d}")
    secondary = (run_secondary(ctx, args.secondary_steps, with_cpu=not args.no_cpu_baseline,
                               threads=args.cpu_threads or cpu_threads()) if args.secondary_steps > 0 else None)
    n2 = sum(r["dim"] == 2 for r in stream)
    npspg = sum(r["kind"] == "PSPG" for r in stream)
    bs, ps = timed, plain_sat
    out = {
        "metric": METRIC,
        "value": bs["steps_per_s"],
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": passes,
        "warmup": warm_passes,
        "ms_per_step": bs["seconds"] * 1e3 / passes,
        "p50_ms": batched_poi["p50_ms"],
        "p99_ms": batched_poi["p99_ms"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64",
        "data": "synthetic",
        "config": {
            "workload": (f"serving stream: {n_req} independent cases (seed 2024), {n2} 2D / {n_req - n2} 3D, "
                         f"{npspg} PSPG dam-break columns / {n_req - npspg} WC sloshing tanks, nodes log-uniform "
                         f"2k-300k, one time step per request; a bench step is one pass of the {n_req}-request "
                         f"stream (each case advances one time step per pass), value = case time steps/s over the "
                         f"timed passes; Poisson arrivals {RATE:.0f}/s for latency"),
            "picard": "Ax_f, minRes 1e-6, maxIter 10",
            "linear_solve": "BiCGSTAB to true relres <= 1e-8 (and ||r|| <= 0.1 minRes), right-preconditioned by "
                            "aggregation multigrid (2D >= 2k nodes; 3D >= 50k FP64 / >= 5k mixed) or block Jacobi",
            "batched": f"continuous batching ({args.slots} slots), PFEM_PREC_MIXED: FP64 BiCGSTAB, multigrid "
                       f"with FP32 level-0 values, FP32 WC state",
            "plain": "one request at a time through the direct calls, FP64 throughout",
            "slots": args.slots, "arena_GiB": args.budget_gb, "stream_cases": n_req,
        },
        "speedup_vs_plain_fp64": bs["pass_steps_per_s"][0] / ps["steps_per_s"],
        "speedup_note": "first timed batched pass (from the initial states) against the plain pass from the same "
                        "states: the same time steps of the same cases",
        "speedup_all_passes_vs_plain_fp64": bs["steps_per_s"] / ps["steps_per_s"],
        "plain_fp64": {"steps_per_s": ps["steps_per_s"], "seconds": ps["seconds"],
                       "poisson_steps_per_s": plain_poi["steps_per_s"], "poisson_p50_ms": plain_poi["p50_ms"],
                       "poisson_p99_ms": plain_poi["p99_ms"], "failed": ps["failed"] + plain_poi["failed"],
                       "max_relres": max(ps["max_relres"], plain_poi["max_relres"]),
                       "krylov_iters": ps["krylov_iters"]},
        "batched_mixed": {"steps_per_s": bs["steps_per_s"], "pass_steps_per_s": bs["pass_steps_per_s"],
                          "poisson_steps_per_s": batched_poi["steps_per_s"],
                          "poisson_p50_ms": batched_poi["p50_ms"], "poisson_p99_ms": batched_poi["p99_ms"],
                          "failed": bs["failed"] + batched_poi["failed"],
                          "max_relres": max(bs["max_relres"], batched_poi["max_relres"]), "ticks": bs["ticks"],
                          "peak_inflight": bs["peak_inflight"], "arena_peak_bytes": bs["arena_peak"],
                          "krylov_iters_pass1": bs["krylov_iters_pass1"]},
        "max_relres": max(bs["max_relres"], batched_poi["max_relres"], ps["max_relres"], plain_poi["max_relres"]),
        "accuracy_vs_fp64": acc,
        "dambreak2d_fp64": db2,
        "dambreak3d_fp64": db,
        "wc3d": wc3d,
        "precision3d": prec3d,
        "secondary_solvers": secondary,
        "roofline": roof,
        "xcd_round_robin": xcd_round_robin,
    }
    if not args.no_cpu_baseline:
        threads = args.cpu_threads or cpu_threads()
        log(f"[bench] CPU baseline with {threads} threads")
        out["cpu_baseline"] = cpu_stream_sample(stream, meshes, states, threads)
        log(f"[bench] CPU baseline (stream sample): {out['cpu_baseline']}")
        out["cpu_configs"] = cpu_configs(threads)
        log(f"[bench] CPU configurations: {out['cpu_configs']}")
    ctx.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
