"""Captured BiCGSTAB iteration chunks (PFEM_GRAPH=1: each chunk of iterations with its V-cycles
captured in a HIP graph and replayed while the engine repeats it) against the same steps launched
kernel by kernel: every step's state and positions equal bit for bit, over several steps whose
replays are separated by the engine's own assemblies, readbacks and synchronisations; block
Jacobi on the 2D dam break (1,001 nodes) and the multigrid on a 2D column of 20,000 nodes."""
import os
import sys

import numpy as np
import pytest

import pfem_amd
from pfem_amd import meshgen as mg

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def run(ctx, mesh, graph, steps, dt, **kw):
    old = os.environ.get("PFEM_GRAPH")
    os.environ["PFEM_GRAPH"] = "1" if graph else "0"
    try:
        c = pfem_amd.Case(ctx, mesh.dim, kind="PSPG", krylov_rtol=1e-10, **kw)
        c.set_mesh(mesh)
        c.set_state(mg.pspg_hydrostatic(mesh))
        out = []
        for _ in range(steps):
            rc, st = c.pspg_step(dt)
            assert rc == 0, st
            out.append((c.get_state(), c.get_positions(), st["picard_iters"], st["krylov_iters"]))
        stats = c.engine_stats()
        c.close()
    finally:
        if old is None:
            del os.environ["PFEM_GRAPH"]
        else:
            os.environ["PFEM_GRAPH"] = old
    return out, stats


@pytest.mark.parametrize("name", ["dambreak2d_jacobi", "column2d_20k_multigrid"])
def test_graph_replays_equal_eager_launches(ctx, name):
    if name == "dambreak2d_jacobi":
        mesh, kw = mg.dambreak_2d(h=0.05), dict(precond="jacobi")
    else:
        mesh, kw = mg.column_2d_nodes(20000), dict(precond="amg")
    say(f"[graph] {name}: N={mesh.n_nodes}")
    eager, s0 = run(ctx, mesh, False, 4, 1e-3, **kw)
    graph, s1 = run(ctx, mesh, True, 4, 1e-3, **kw)
    say(f"  eager {s0}  graph {s1}")
    assert s0["graph_launches"] == 0
    assert s1["graph_captures"] >= 1 and s1["graph_launches"] > s1["graph_captures"]
    for k, ((qa, pa, ia, ka), (qb, pb, ib, kb)) in enumerate(zip(eager, graph)):
        assert (ia, ka) == (ib, kb), k
        assert np.array_equal(qa, qb), k
        assert np.array_equal(pa, pb), k
