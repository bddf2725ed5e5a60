"""Parity at the configurations' real sizes.

* 3D dam break, 262,701 nodes / 1,500,000 tets: the assembled system with its boundary
  conditions against the oracle (CSR pattern bit-exact, values, rhs), then one Picard step
  whose converged state satisfies the oracle's own residual test (the Ax_f residual of the
  oracle's system assembled at the step's final positions is <= minRes).
* Random-walk column field, 1,067,788 nodes: the mixed-precision solve (FP32 inner BiCGSTAB
  inside FP64 refinement) to relres 1e-10, measured on the host with the exported FP64
  system; the exported rows of ~2,000 sampled nodes against the oracle's assembly of the
  sub-mesh of their elements (an element-complete neighbourhood, so those rows are whole).
* 64 requests of the seeded serving stream (2k-300k nodes, both kinds, both dimensions):
  each result served in the mix equals the same case stepped alone, bit for bit.

The oracle's assembly of the full dam break takes ~30 s and ~18 GB of host memory; the
full random-walk field would need ~75 GB, hence the sub-mesh rows there.
"""
import os
import sys
import time

import numpy as np
import pytest

import pfem_amd
from oracle import oracle as O
from pfem_amd import meshgen as mg

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def value_errors(rp, ref, got):
    """(largest |got - ref| relative to each row's largest magnitude, largest per-value relative
    error over the entries of at least 1e-3 of their row's largest magnitude)."""
    row = np.repeat(np.arange(len(rp) - 1), np.diff(rp))
    rmax = np.zeros(len(rp) - 1)
    np.maximum.at(rmax, row, np.abs(ref))
    scale = rmax[row]
    d = np.abs(got - ref)
    nz = scale > 0
    row_rel = float((d[nz] / scale[nz]).max()) if nz.any() else 0.0
    assert np.all(d[~nz] == 0)
    big = np.abs(ref) >= 1e-3 * scale
    big &= nz
    val_rel = float((d[big] / np.abs(ref[big])).max()) if big.any() else 0.0
    return row_rel, val_rel


def block_rel_err(d, N, ref, got):
    worst = 0.0
    for s in range(d + 1):
        r, g = ref[s * N:(s + 1) * N], got[s * N:(s + 1) * N]
        scale = np.abs(r).max()
        worst = max(worst, np.abs(g - r).max() / scale if scale > 0 else np.abs(g).max())
    return worst


def test_dambreak3d_fullsize_assembly_and_picard_step(ctx):
    mesh = mg.dambreak_3d(h=0.01, col=(0.5, 0.5, 1.0), seed=0)
    d, N = 3, mesh.n_nodes
    assert N == 262701 and mesh.n_elems == 1500000
    dt = 5e-4
    rng = np.random.default_rng(21)
    q0 = mg.pspg_hydrostatic(mesh)
    q_iter = q0.copy()
    moving = (mesh.flags & mg.BOUND) == 0
    q_iter[:d * N] = (rng.normal(scale=0.05, size=(d, N)) * moving).reshape(-1)
    dv = np.zeros(d * N)
    say(f"[C2 full] N={N} E={mesh.n_elems}: oracle assembly")
    t0 = time.time()
    ref = O.pspg_assemble(mesh, q0, q_iter, dir_vals=dv, dt=dt, apply_bc=True)
    say(f"  oracle {time.time() - t0:.1f} s, nnz {ref.row_ptr[-1]}")
    c = pfem_amd.Case(ctx, d, kind="PSPG", residual="Ax_f", min_res=1e-6, max_iter=10, krylov_rtol=1e-9)
    c.set_mesh(mesh)
    c.assemble(dt, q_prev=q0, q_iter=q_iter)
    rp, ci, va = c.csr()
    np.testing.assert_array_equal(rp, ref.row_ptr)
    np.testing.assert_array_equal(ci, ref.col_idx)
    row_rel, val_rel = value_errors(rp, ref.vals, va)
    b_rel = block_rel_err(d, N, ref.b, c.rhs())
    say(f"  values: row-relative {row_rel:.2e}, per-value (entries >= 1e-3 row max) {val_rel:.2e}; b {b_rel:.2e}")
    assert row_rel <= 1e-12
    assert val_rel <= 1e-11
    assert b_rel <= 1e-12
    del ref, va, ci
    # one Picard step from the hydrostatic state; the converged iterate passes the oracle's
    # residual test on the oracle's system at the final positions (PSPG.inl:298-301,368)
    c.set_state(q0)
    c.set_positions(mesh.pos)
    rc, st = c.pspg_step(dt)
    say(f"  step: rc={rc} {st}")
    assert rc == 0
    q1 = c.get_state()
    pos1 = c.get_positions()
    moved = mg.PfemMesh(d, pos1, mesh.conn, mesh.flags, mesh.h)
    t0 = time.time()
    S1 = O.pspg_assemble(moved, q0, q1, dir_vals=dv, dt=dt, apply_bc=True)
    res = float(np.linalg.norm(S1.to_scipy() @ q1 - S1.b))
    say(f"  oracle residual at the final positions {res:.3e} ({time.time() - t0:.1f} s); GPU {st['residual']:.3e}")
    assert res <= 1e-6
    assert res == pytest.approx(st["residual"], rel=1e-3, abs=1e-9)
    # positions moved from the saved ones by dt v (fixed wall nodes stay)
    fixed = (mesh.flags & mg.FIXED) != 0
    np.testing.assert_array_equal(pos1[:, fixed], mesh.pos[:, fixed])
    np.testing.assert_allclose(pos1[:, ~fixed], mesh.pos[:, ~fixed] + dt * q1.reshape(d + 1, N)[:d, ~fixed],
                               rtol=0, atol=1e-15)
    c.close()


def sub_mesh(mesh, nodes):
    """The elements touching `nodes` and their nodes, renumbered in the original order."""
    touch = np.isin(mesh.conn, nodes).any(axis=1)
    conn = mesh.conn[touch]
    keep = np.unique(conn)
    new = np.full(mesh.n_nodes, -1, np.int64)
    new[keep] = np.arange(keep.size)
    sm = mg.PfemMesh(mesh.dim, mesh.pos[:, keep].copy(), new[conn].astype(np.int32), mesh.flags[keep].copy(),
                     mesh.h)
    return sm, keep, new


def test_randomwalk_fullsize_mixed_ir_solve_and_sampled_rows(ctx):
    mesh = mg.randomwalk_3d(h=0.01, ncol=100, mean_cells=100, seed=0)
    d, N = 3, mesh.n_nodes
    say(f"[C4 full] N={N} E={mesh.n_elems}")
    assert N > 1_000_000
    dt = 5e-4
    q0 = mg.pspg_hydrostatic(mesh)
    rng = np.random.default_rng(3)
    dv = rng.normal(scale=0.01, size=d * N)
    c = pfem_amd.Case(ctx, d, kind="PSPG", precision="mixed_ir", precond="jacobi", krylov_rtol=1e-10,
                      krylov_max_iter=40000, inner_rtol=3e-2, inner_max_iter=200)
    c.set_mesh(mesh)
    c.set_dirichlet(dv)
    c.assemble(dt, q_prev=q0, q_iter=q0)
    t0 = time.time()
    rc, x, st = c.solve(x0=q0)
    say(f"  mixed-IR solve rc={rc} {st} ({time.time() - t0:.2f} s)")
    assert rc == 0
    rp, ci, va = c.csr()
    b = c.rhs()
    import scipy.sparse as sp

    A = sp.csr_matrix((va, ci, rp), shape=(len(b), len(b)))
    relres = float(np.linalg.norm(b - A @ x) / np.linalg.norm(b))
    say(f"  host relres {relres:.3e}")
    assert relres <= 1e-10
    # ~2,000 sampled nodes: their rows against the oracle's assembly of their elements
    sample = np.unique(rng.choice(N, size=2000, replace=False))
    sample = sample[sample != 0]
    sm, keep, new = sub_mesh(mesh, sample)
    n_sub = sm.n_nodes
    S = O.pspg_assemble(sm, q0.reshape(d + 1, N)[:, keep].reshape(-1), q0.reshape(d + 1, N)[:, keep].reshape(-1),
                        dir_vals=dv.reshape(d, N)[:, keep].reshape(-1), dt=dt, apply_bc=True)
    worst_row, worst_b = 0.0, 0.0
    for s in range(d + 1):
        rows_g = sample + s * N
        rows_o = new[sample] + s * n_sub
        for rg, ro in zip(rows_g, rows_o):
            cg = ci[rp[rg]:rp[rg + 1]]
            vg = va[rp[rg]:rp[rg + 1]]
            co = S.col_idx[S.row_ptr[ro]:S.row_ptr[ro + 1]]
            vo = S.vals[S.row_ptr[ro]:S.row_ptr[ro + 1]]
            mapped = new[cg % N] + (cg // N) * n_sub
            np.testing.assert_array_equal(mapped, co)
            scale = np.abs(vo).max()
            worst_row = max(worst_row, np.abs(vg - vo).max() / scale if scale > 0 else np.abs(vg).max())
        bo = S.b[rows_o]
        bg = b[rows_g]
        worst_b = max(worst_b, np.abs(bg - bo).max() / max(np.abs(bo).max(), 1e-300))
    say(f"  sampled rows: {sample.size} nodes, row-relative {worst_row:.2e}, b {worst_b:.2e}")
    assert worst_row <= 1e-12
    assert worst_b <= 1e-12
    c.close()


def test_serving_stream_slice_alone_equals_in_the_mix(ctx):
    sys.path.insert(0, ROOT)
    import bench

    stream = bench.build_stream(512)
    # 64 requests spanning the stream: the largest of each kind/dimension, then every 8th
    pick = set()
    for kind in ("PSPG", "WC"):
        for dim in (2, 3):
            grp = [r for r in stream if r["kind"] == kind and r["dim"] == dim]
            pick.add(max(grp, key=lambda r: r["target"])["i"])
    for r in stream[::8]:
        if len(pick) == 64:
            break
        pick.add(r["i"])
    sl = [r for r in stream if r["i"] in pick]
    sizes = [r["target"] for r in sl]
    say(f"[C5 slice] {len(sl)} requests, nodes {min(sizes)}-{max(sizes)}, "
        f"{sum(r['kind'] == 'PSPG' for r in sl)} PSPG, {sum(r['dim'] == 3 for r in sl)} 3D")
    assert len(sl) >= 64 and max(sizes) >= 200_000 and min(sizes) <= 5_000
    meshes = {r["i"]: bench.make_mesh(r) for r in sl}
    states = {r["i"]: bench.initial_state(r, meshes[r["i"]]) for r in sl}

    def fresh(r):
        c = bench.make_case(ctx, r, meshes[r["i"]], "batched")
        c.set_state(states[r["i"]])
        return c

    for r in sl:
        r["dt"] = (1e-3 if r["dim"] == 2 else 5e-4) if r["kind"] == "PSPG" else 2e-6
    alone = {}
    for r in sl:
        c = fresh(r)
        if r["kind"] == "PSPG":
            rc, st = c.pspg_step(r["dt"])
            assert rc == 0, (r, st)
        else:
            c.wc_step(r["dt"])
        alone[r["i"]] = (c.get_state(), c.get_positions().copy())
        c.close()
    cases = {r["i"]: fresh(r) for r in sl}
    srv = pfem_amd.Server(ctx, hbm_budget=64 << 30, max_inflight=64, wc_security_coeff=0.5, wc_max_dt=1e-3)
    t0 = pfem_amd.now_ns()
    for k, r in enumerate(sl):  # arrivals 0.5 ms apart: most join a running batch
        srv.submit(cases[r["i"]], r["dt"], tag=r["i"], arrival_ns=t0 + k * 500_000)
    got = 0
    deadline = time.time() + 120
    while got < len(sl):
        assert time.time() < deadline, "server did not finish"
        for res in srv.poll(timeout_ms=200):
            assert res["status"] == 0, res
            got += 1
    st = srv.stats()
    say(f"  server stats {st}")
    srv.close()
    for r in sl:
        q, x = cases[r["i"]].get_state(), cases[r["i"]].get_positions()
        q0, x0 = alone[r["i"]]
        assert np.array_equal(q, q0) and np.array_equal(x, x0), f"request {r['i']} ({r['kind']} {r['dim']}D)"
        cases[r["i"]].close()
