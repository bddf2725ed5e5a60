"""GPU parity on edge-case meshes against the CPU oracle, through the C-ABI: a single
triangle and a single tetrahedron (one ELL slice with 3 or 4 live rows), each with an
isolated node that the library flags free; the unit square in 2 triangles and the unit cube
in 6 Kuhn tetrahedra; and a ragged field of random-walk columns (uneven node degrees).

Bars as in the other parity tests: CSR pattern bit-exact, values and b within 1e-12 of the
oracle relative to each row's / field block's largest magnitude, solutions to relres 1e-9
on the oracle's own system, WC steps within 1e-10 (FP64 state)."""
import itertools
import sys

import numpy as np
import pytest

import pfem_amd
from oracle import oracle as O
from pfem_amd import meshgen as mg

pytestmark = pytest.mark.gpu

K0, K0P, RS, MU = 2.2e6, 7.0, 1000.0, 1e-3


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def oriented(pos, conn):
    conn = np.array(conn, dtype=np.int32)
    for e in range(len(conn)):
        P = pos[:, conn[e]]
        if np.linalg.det(P[:, 1:] - P[:, :1]) < 0:
            conn[e, [0, 1]] = conn[e, [1, 0]]
    return conn


def tiny_meshes():
    out = {}
    # one triangle / one tetrahedron, plus an isolated node (free: it belongs to no element)
    pos = np.array([[0.0, 1.0, 0.0, 2.0], [0.0, 0.0, 1.0, 2.0]])
    fl = np.array([mg.BOUND | mg.DIRICHLET, mg.BOUND, 0, 0], np.uint8)
    out["one_triangle"] = mg.PfemMesh(2, pos, oriented(pos, [[0, 1, 2]]), fl, 1.0, name="one_triangle")
    pos = np.array([[0.0, 1.0, 0.0, 0.0, 2.0], [0.0, 0.0, 1.0, 0.0, 2.0], [0.0, 0.0, 0.0, 1.0, 2.0]])
    fl = np.array([mg.BOUND | mg.DIRICHLET, mg.BOUND, 0, 0, 0], np.uint8)
    out["one_tet"] = mg.PfemMesh(3, pos, oriented(pos, [[0, 1, 2, 3]]), fl, 1.0, name="one_tet")
    # unit square (2 triangles) and Kuhn cube (6 tetrahedra)
    pos = np.array([[0.0, 1.0, 0.0, 1.0], [0.0, 0.0, 1.0, 1.0]])
    fl = np.array([mg.BOUND | mg.DIRICHLET, mg.BOUND, 0, 0], np.uint8)
    out["unit_square"] = mg.PfemMesh(2, pos, oriented(pos, [[0, 1, 3], [0, 3, 2]]), fl, 1.0, name="unit_square")
    pts = [(x, y, z) for z in (0, 1) for y in (0, 1) for x in (0, 1)]
    pos = np.array(pts, dtype=float).T
    conn = []
    for perm in itertools.permutations(range(3)):
        v, path = 0, [0]
        for ax in perm:
            v |= 1 << ax
            path.append(v)
        conn.append(path)
    fl = np.zeros(8, np.uint8)
    fl[[0, 1, 2, 4]] = mg.BOUND
    fl[0] |= mg.DIRICHLET
    out["kuhn_cube"] = mg.PfemMesh(3, pos, oriented(pos, conn), fl, 1.0, name="kuhn_cube")
    return out


def ragged_mesh():
    return mg.randomwalk_3d(h=0.05, ncol=8, mean_cells=20, seed=1)


def all_meshes():
    m = tiny_meshes()
    m["ragged_3d"] = ragged_mesh()
    return m


def rand_pspg_state(mesh, seed):
    rng = np.random.default_rng(seed)
    d, N = mesh.dim, mesh.n_nodes
    q = np.zeros((d + 1) * N)
    q[:d * N] = rng.normal(scale=0.2, size=d * N)
    q[d * N:] = rng.uniform(0.0, 2000.0, size=N)
    return q


def row_rel_err(rp, ref, got):
    worst = 0.0
    for r in range(len(rp) - 1):
        a, b = rp[r], rp[r + 1]
        scale = np.abs(ref[a:b]).max() if b > a else 0.0
        if scale == 0:
            assert np.all(got[a:b] == 0)
            continue
        worst = max(worst, np.abs(got[a:b] - ref[a:b]).max() / scale)
    return worst


def block_rel_err(d, N, ref, got):
    worst = 0.0
    for s in range(d + 1):
        r = ref[s * N:(s + 1) * N]
        g = got[s * N:(s + 1) * N]
        scale = np.abs(r).max()
        worst = max(worst, np.abs(g - r).max() / scale if scale > 0 else np.abs(g).max())
    return worst


@pytest.mark.parametrize("name", ["one_triangle", "one_tet", "unit_square", "kuhn_cube", "ragged_3d"])
@pytest.mark.parametrize("with_bc", [False, True])
def test_edge_assembly_parity(ctx, name, with_bc):
    mesh = all_meshes()[name]
    d, N = mesh.dim, mesh.n_nodes
    say(f"[edge assembly] {name} bc={with_bc} N={N} E={mesh.n_elems}")
    qprev, qiter = rand_pspg_state(mesh, 1), rand_pspg_state(mesh, 2)
    rng = np.random.default_rng(5)
    dv = rng.normal(scale=0.5, size=d * N) if with_bc else np.zeros(d * N)
    dt = 1e-3 if d == 2 else 5e-4
    ref = O.pspg_assemble(mesh, qprev, qiter, dir_vals=dv, dt=dt, apply_bc=True)
    c = pfem_amd.Case(ctx, d, kind="PSPG")
    c.set_mesh(mesh)
    c.set_dirichlet(dv)
    c.assemble(dt, q_prev=qprev, q_iter=qiter)
    rp, ci, va = c.csr()
    np.testing.assert_array_equal(rp, ref.row_ptr)
    np.testing.assert_array_equal(ci, ref.col_idx)
    assert row_rel_err(rp, ref.vals, va) <= 1e-12
    assert block_rel_err(d, N, ref.b, c.rhs()) <= 1e-12
    c.close()


@pytest.mark.parametrize("precision", ["fp64", "mixed_ir"])
@pytest.mark.parametrize("name", ["one_triangle", "one_tet", "kuhn_cube", "ragged_3d"])
def test_edge_solve_relres(ctx, name, precision):
    mesh = all_meshes()[name]
    d = mesh.dim
    say(f"[edge solve] {name} {precision}")
    q = rand_pspg_state(mesh, 4)
    dt = 1e-3 if d == 2 else 5e-4
    ref = O.pspg_assemble(mesh, q, q, dt=dt)
    c = pfem_amd.Case(ctx, d, kind="PSPG", precision=precision, krylov_rtol=1e-10)
    c.set_mesh(mesh)
    c.assemble(dt, q_prev=q, q_iter=q)
    rc, x, st = c.solve(x0=np.zeros_like(q))
    assert rc == 0, st
    A = ref.to_scipy()
    relres = np.linalg.norm(ref.b - A @ x) / np.linalg.norm(ref.b)
    say(f"  iters={st['krylov_iters']} relres={relres:.2e}")
    assert relres <= 1e-9, (relres, st)
    c.close()


@pytest.mark.parametrize("name", ["one_triangle", "one_tet", "unit_square", "kuhn_cube"])
def test_edge_wc_step_parity(ctx, name):
    mesh = tiny_meshes()[name]
    d, N = mesh.dim, mesh.n_nodes
    say(f"[edge wc] {name} N={N}")
    rng = np.random.default_rng(7)
    v = rng.normal(scale=0.1, size=d * N)
    p = rng.uniform(0.0, 5000.0, size=N)
    rho = RS * (K0P / K0 * p + 1.0) ** (1.0 / K0P)
    a = np.zeros(d * N)
    c = pfem_amd.Case(ctx, d, kind="WC", mu=MU, K0=K0, K0p=K0P, rho_star=RS, wc_state="fp64")
    c.set_mesh(mesh)
    c.set_state(np.concatenate([v, p, rho, a]))
    pos_o = mesh.pos.reshape(-1).copy()
    vo, po, ro, ao = v.copy(), p.copy(), rho.copy(), a.copy()
    for step in range(2):
        # max_dt far above the CFL bound of these unit-size elements, so the bound is compared
        dt_o = O.wc_next_dt(mesh, pos_o, vo, po, ro, mu=MU, K0=K0, K0p=K0P, security_coeff=0.5, max_dt=1.0)
        dt_g = c.wc_next_dt(0.5, 1.0)
        assert dt_o < 1.0
        assert dt_g == pytest.approx(dt_o, rel=1e-12)
        O.wc_step(mesh, pos_o, vo, po, ro, ao, dt_o, mu=MU, K0=K0, K0p=K0P, rho_star=RS)
        c.wc_step(dt_o)
        s = c.get_state()
        vg, pg, rg, ag = s[:d * N], s[d * N:(d + 1) * N], s[(d + 1) * N:(d + 2) * N], s[(d + 2) * N:]
        for ref, got in ((vo, vg), (po, pg), (ro, rg)):
            scale = max(np.abs(ref).max(), 1.0)
            assert np.abs(got - ref).max() <= 1e-10 * scale
        assert np.abs(ag - ao).max() <= 1e-10 * max(np.abs(ao).max(), 9.81)
        np.testing.assert_allclose(c.get_positions().reshape(-1), pos_o, rtol=0, atol=1e-12)
    c.close()


def test_edge_picard_step_parity_ragged(ctx):
    """Two Picard time steps on the ragged column field from its hydrostatic state, against
    the oracle's Picard loop with a tight Krylov stand-in."""
    mesh = ragged_mesh()
    d, N = mesh.dim, mesh.n_nodes
    say(f"[edge picard] ragged N={N}")
    dt = 5e-4
    q0 = mg.pspg_hydrostatic(mesh)
    pos_o = mesh.pos.reshape(-1).copy()
    q_o = q0.copy()
    c = pfem_amd.MomContEqIncompNewton(ctx, d, krylov_rtol=1e-11)
    c.set_mesh(mesh)
    c.set_state(q0)
    for step in range(2):
        st_o, info = O.pspg_step(mesh, pos_o, q_o, dt=dt, solver_kind=2, rtol=1e-13)
        rc, st = c.pspg_step(dt)
        say(f"  step {step}: oracle iters={info['picard_iters']} gpu iters={st['picard_iters']}")
        assert st_o == 0 and rc == 0
        assert abs(st["picard_iters"] - info["picard_iters"]) <= 1
        assert block_rel_err(d, N, q_o, c.get_state()) <= 1e-6
        np.testing.assert_allclose(c.get_positions().reshape(-1), pos_o, rtol=0, atol=1e-9 * np.abs(pos_o).max())
    c.close()
