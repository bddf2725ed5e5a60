"""Batched and served steps of many independent cases, through the C-ABI.

The bar: a case stepped inside a batch, or inside the continuous-batching server with other
requests arriving while it runs, gives the same result as the same case stepped alone, bit
for bit: the PSPG assembly and the WC node sums gather their element contributions in a
fixed order with no atomics, every reduction is per case in a fixed order, and a case's
launches do not depend on the batch.
"""
import sys
import time

import numpy as np
import pytest

import pfem_amd
from pfem_amd import meshgen as mg

pytestmark = pytest.mark.gpu

K0, K0P, RS = 2.2e6, 7.0, 1000.0


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def specs():
    """(kind, dim, precision / state, mesh, preconditioner) of a small mixed population."""
    return [
        ("PSPG", 2, "fp64", mg.dambreak_2d(), "jacobi"),
        ("PSPG", 2, "mixed", mg.column_2d_nodes(900, h=0.02, seed=4), "jacobi"),
        ("PSPG", 3, "fp64", mg.dambreak_3d(h=0.05, col=(0.5, 0.5, 1.0), seed=1), "amg"),
        ("PSPG", 3, "mixed", mg.column_3d_nodes(3000, h=0.02, seed=5), "jacobi"),
        ("WC", 2, "fp64", mg.tank_2d(h=0.05, size=(1.0, 0.5), fill=0.6, seed=2), None),
        ("WC", 3, "fp32", mg.tank_3d(h=0.05, size=(0.4, 0.4, 0.3), fill=0.7, seed=3), None),
        ("PSPG", 2, "fp64", mg.column_2d_nodes(2500, h=0.01, seed=6), "amg"),
        ("WC", 3, "fp64", mg.tank_nodes(3, 4000, h=0.02, seed=7), None),
        ("PSPG", 2, "mixed", mg.column_2d_nodes(6000, h=0.01, seed=8), "amg"),
        ("PSPG", 3, "mixed_ir", mg.column_3d_nodes(3000, h=0.02, seed=9), "jacobi"),
        ("PSPG", 3, "mixed", mg.column_3d_nodes(8000, h=0.02, seed=10), "amg"),
        ("PSPG", 3, "fp64", mg.randomwalk_3d(h=0.05, ncol=8, mean_cells=20, seed=1), "jacobi"),
        ("WC", 3, "fp64", mg.randomwalk_3d(h=0.05, ncol=8, mean_cells=20, seed=1), None),
    ]


def make(ctx, spec):
    kind, dim, prec, mesh, precond = spec
    if kind == "PSPG":
        c = pfem_amd.Case(ctx, dim, kind="PSPG", precision=prec, krylov_rtol=1e-9, precond=precond)
        c.set_mesh(mesh)
        c.set_state(mg.pspg_hydrostatic(mesh))
    else:
        c = pfem_amd.Case(ctx, dim, kind="WC", K0=K0, K0p=K0P, rho_star=RS, wc_state=prec)
        c.set_mesh(mesh)
        v, p, rho, a = mg.wc_tait_hydrostatic(mesh, K0=K0, K0p=K0P, rho_star=RS)
        c.set_state(np.concatenate([v.reshape(-1), p, rho, a.reshape(-1)]))
    return c


def dt_of(spec):
    return (1e-3 if spec[1] == 2 else 5e-4) if spec[0] == "PSPG" else 2e-6


def alone(ctx, spec, steps=1):
    c = make(ctx, spec)
    for _ in range(steps):
        if spec[0] == "PSPG":
            rc, st = c.pspg_step(dt_of(spec))
            assert rc == 0, st
        else:
            c.wc_step(dt_of(spec))
    out = (c.get_state(), c.get_positions().copy())
    c.close()
    return out


def same(spec, ref, got):
    (q0, x0), (q1, x1) = ref, got
    return np.array_equal(q0, q1) and np.array_equal(x0, x1)


def test_batch_step_equals_alone(ctx):
    sp = specs()
    refs = [alone(ctx, s) for s in sp]
    cases = [make(ctx, s) for s in sp]
    say(f"[batch] {len(cases)} cases")
    status, stats = pfem_amd.batch_step(ctx, cases, [dt_of(s) for s in sp])
    assert list(status) == [0] * len(sp), stats
    for i, (s, c) in enumerate(zip(sp, cases)):
        got = (c.get_state(), c.get_positions().copy())
        assert same(s, refs[i], got), f"case {i} {s[:3]} differs from its alone run"
        c.close()


def test_server_results_equal_alone_with_staggered_arrivals(ctx):
    sp = specs()
    refs = [alone(ctx, s, steps=2) for s in sp]
    cases = [make(ctx, s) for s in sp]
    srv = pfem_amd.Server(ctx, hbm_budget=1 << 30, max_inflight=4)
    say("[server] staggered arrivals, 2 steps per case")
    t0 = pfem_amd.now_ns()
    pending = {}
    # first step of every case, arrivals 2 ms apart (most arrive while others run)
    for i, (s, c) in enumerate(zip(sp, cases)):
        srv.submit(c, dt_of(s), tag=i, arrival_ns=t0 + i * 2_000_000)
        pending[i] = 1
    done_steps = {i: 0 for i in range(len(sp))}
    results = []
    deadline = time.time() + 60
    while any(v < 2 for v in done_steps.values()):
        assert time.time() < deadline, "server did not finish"
        for r in srv.poll(timeout_ms=200):
            i = r["tag"]
            assert r["status"] == 0, r
            assert r["t_arrival_ns"] <= r["t_admit_ns"] <= r["t_done_ns"]
            results.append(r)
            done_steps[i] += 1
            if done_steps[i] == 1:  # second step submitted as soon as the first returns
                srv.submit(cases[i], dt_of(sp[i]), tag=i)
    st = srv.stats()
    say(f"  stats {st}")
    assert st["completed"] == 2 * len(sp) and st["peak_inflight"] <= 4
    for i, (s, c) in enumerate(zip(sp, cases)):
        got = (c.get_state(), c.get_positions().copy())
        assert same(s, refs[i], got), f"case {i} {s[:3]} differs from its alone run"
    wc = [r for r in results if r["kind"] == 1]
    assert all(r["next_dt"] > 0 for r in wc)
    srv.close()
    for c in cases:
        c.close()


def test_server_admission_by_workspace_footprint(ctx):
    """An arena that holds the largest workspace once (but not two of them) still serves
    every request, never exceeds its size, and gives the alone results."""
    sp = [s for s in specs() if s[0] == "PSPG"]
    refs = [alone(ctx, s) for s in sp]
    cases = [make(ctx, s) for s in sp]
    ws = [c.workspace_bytes() for c in cases]
    budget = int(max(ws) * 1.5)
    srv = pfem_amd.Server(ctx, hbm_budget=budget, max_inflight=16)
    say(f"[admission] workspaces {ws} budget {budget}")
    for i, (s, c) in enumerate(zip(sp, cases)):
        srv.submit(c, dt_of(s), tag=i)
    got = 0
    deadline = time.time() + 60
    while got < len(sp):
        assert time.time() < deadline, "server did not finish"
        for r in srv.poll(timeout_ms=200):
            assert r["status"] == 0, r
            got += 1
    st = srv.stats()
    say(f"  stats {st}")
    assert st["arena_peak"] <= st["arena_bytes"]
    assert st["arena_used"] == 0
    for i, (s, c) in enumerate(zip(sp, cases)):
        assert same(s, refs[i], (c.get_state(), c.get_positions().copy())), f"case {i}"
    srv.close()
    for c in cases:
        c.close()


def test_server_rejects_oversized_and_busy(ctx):
    sp = specs()
    big = make(ctx, sp[2])
    srv = pfem_amd.Server(ctx, hbm_budget=big.workspace_bytes() // 2, max_inflight=2)
    with pytest.raises(pfem_amd.PfemError) as e:
        srv.submit(big, 5e-4, tag=0)
    assert e.value.code == -2  # PFEM_E_OOM
    small = make(ctx, sp[0])
    srv2 = pfem_amd.Server(ctx, hbm_budget=1 << 28, max_inflight=2)
    srv2.submit(small, 1e-3, tag=1, arrival_ns=pfem_amd.now_ns() + 200_000_000)
    with pytest.raises(pfem_amd.PfemError) as e:
        srv2.submit(small, 1e-3, tag=2)
    assert e.value.code == -5  # PFEM_E_STATE: one pending request per case
    with pytest.raises(pfem_amd.PfemError):
        small.get_state()  # held by the server until its result is polled
    res = []
    deadline = time.time() + 30
    while not res:
        assert time.time() < deadline
        res = srv2.poll(timeout_ms=200)
    assert res[0]["status"] == 0 and res[0]["t_admit_ns"] >= res[0]["t_arrival_ns"]
    srv.close()
    srv2.close()
    big.close()
    small.close()
