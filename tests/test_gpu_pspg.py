"""GPU parity of the PSPG path (K0 pattern, K1 assembly, K2 BC, K3/K5 solve, Picard step)
against the CPU oracle, through the C-ABI.

Bars: CSR pattern (row_ptr, col_idx) bit-exact; assembled values within 1e-12 of the
oracle relative to each row's largest magnitude (the GPU gathers each entry's element
contributions in element order, Eigen's triplet order, but evaluates the element terms
with its own operation order);
b within 1e-12 relative to its field block's largest magnitude; solutions to the stated
relative residual on the oracle's own system.
"""
import sys

import numpy as np
import pytest

import pfem_amd
from oracle import oracle as O
from pfem_amd import meshgen as mg

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def small_meshes():
    return {
        "2d_dambreak": mg.dambreak_2d(),  # has free wall nodes
        "2d_column": mg.column_2d_nodes(400, h=0.02, seed=3),
        "3d_column": mg.dambreak_3d(h=0.05, col=(0.5, 0.5, 1.0), seed=1),
    }


def rand_state(mesh, seed):
    rng = np.random.default_rng(seed)
    q = mg.pspg_hydrostatic(mesh)
    d = mesh.dim
    N = mesh.n_nodes
    q[:d * N] = rng.normal(scale=0.2, size=d * N)
    return q


def row_rel_err(rp, ref, got):
    worst = 0.0
    for r in range(len(rp) - 1):
        a, b = rp[r], rp[r + 1]
        scale = np.abs(ref[a:b]).max()
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


def make_case(ctx, mesh, **kw):
    c = pfem_amd.Case(ctx, mesh.dim, kind="PSPG", **kw)
    c.set_mesh(mesh)
    return c


@pytest.mark.parametrize("name", ["2d_dambreak", "2d_column", "3d_column"])
@pytest.mark.parametrize("with_bc", [False, True])
def test_assembly_parity(ctx, name, with_bc):
    mesh = small_meshes()[name]
    say(f"[assembly] {name} bc={with_bc} N={mesh.n_nodes} E={mesh.n_elems}")
    d, N = mesh.dim, mesh.n_nodes
    qprev = rand_state(mesh, 1)
    qiter = rand_state(mesh, 2)
    rng = np.random.default_rng(5)
    dv = rng.normal(scale=0.5, size=d * N) if with_bc else np.zeros(d * N)
    dt = 1e-3 if d == 2 else 5e-4
    ref = O.pspg_assemble(mesh, qprev, qiter, dir_vals=dv, dt=dt, apply_bc=True)
    c = make_case(ctx, mesh)
    c.set_dirichlet(dv)
    c.assemble(dt, q_prev=qprev, q_iter=qiter)
    rp, ci, va = c.csr()
    np.testing.assert_array_equal(rp, ref.row_ptr)
    np.testing.assert_array_equal(ci, ref.col_idx)
    assert row_rel_err(rp, ref.vals, va) <= 1e-12
    assert block_rel_err(d, N, ref.b, c.rhs()) <= 1e-12
    c.close()


def test_pattern_is_reproducible(ctx):
    mesh = small_meshes()["3d_column"]
    c1 = make_case(ctx, mesh)
    c2 = make_case(ctx, mesh)
    for c in (c1, c2):
        c.assemble(5e-4, q_prev=rand_state(mesh, 1))
    a, b = c1.csr(), c2.csr()
    np.testing.assert_array_equal(a[0], b[0])
    np.testing.assert_array_equal(a[1], b[1])


@pytest.mark.parametrize("name", ["2d_dambreak", "3d_column"])
def test_spmv_matches_scipy(ctx, name):
    mesh = small_meshes()[name]
    say(f"[spmv] {name}")
    c = make_case(ctx, mesh)
    q = rand_state(mesh, 3)
    c.assemble(1e-3, q_prev=q)
    rp, ci, va = c.csr()
    import scipy.sparse as sp

    A = sp.csr_matrix((va, ci, rp))
    x = np.random.default_rng(9).normal(size=A.shape[0])
    y = c.spmv(x)
    ref = A @ x
    assert np.abs(y - ref).max() <= 1e-12 * np.abs(A).sum(axis=1).max() * np.abs(x).max()


@pytest.mark.parametrize("precond", ["jacobi", "amg"])
@pytest.mark.parametrize("precision", ["fp64", "mixed", "mixed_ir"])
@pytest.mark.parametrize("name", ["2d_dambreak", "3d_column"])
def test_solve_relres_on_oracle_system(ctx, name, precision, precond):
    mesh = small_meshes()[name]
    say(f"[solve] {name} {precision} {precond}")
    d, N = mesh.dim, mesh.n_nodes
    q = rand_state(mesh, 4)
    dt = 1e-3 if d == 2 else 5e-4
    ref = O.pspg_assemble(mesh, q, q, dt=dt)
    c = make_case(ctx, mesh, precision=precision, krylov_rtol=1e-10, precond=precond)
    c.assemble(dt, q_prev=q, q_iter=q)
    rc, x, st = c.solve(x0=np.zeros_like(q))
    assert rc == 0, st
    A = ref.to_scipy()
    relres = np.linalg.norm(ref.b - A @ x) / np.linalg.norm(ref.b)
    assert relres <= 1e-9, (relres, st)
    import scipy.sparse.linalg as spla

    xs = spla.spsolve(A.tocsc(), ref.b)
    assert np.linalg.norm(x - xs) / np.linalg.norm(xs) <= 1e-6


@pytest.mark.parametrize("precision", ["fp64", "mixed", "mixed_all_fp32"])
@pytest.mark.parametrize("dim", [2, 3])
def test_amg_cuts_iterations(ctx, dim, precision, monkeypatch):
    """On a mesh large enough for a multi-level hierarchy, the multigrid-preconditioned solve
    reaches the same true relative residual on the oracle's system in a fraction of the
    block-Jacobi iterations (mixed: FP32 level-0 values under the FP64 hierarchy;
    mixed_all_fp32: PFEM_AMG_FP32=1, a 3D hierarchy wholly in FP32, 2D as mixed)."""
    if precision == "mixed_all_fp32":
        monkeypatch.setenv("PFEM_AMG_FP32", "1")
        precision = "mixed"
    mesh = mg.column_2d_nodes(20000, h=0.01, seed=2) if dim == 2 else mg.column_3d_nodes(30000, h=0.01, seed=2)
    say(f"[amg] dim={dim} {precision} N={mesh.n_nodes}")
    q = mg.pspg_hydrostatic(mesh)
    dt = 1e-3 if dim == 2 else 5e-4
    ref = O.pspg_assemble(mesh, q, q, dt=dt)
    A = ref.to_scipy()
    its = {}
    for precond in ("jacobi", "amg"):
        c = make_case(ctx, mesh, precision=precision, krylov_rtol=1e-8, precond=precond)
        c.assemble(dt, q_prev=q, q_iter=q)
        rc, x, st = c.solve(x0=np.zeros_like(q))
        relres = np.linalg.norm(ref.b - A @ x) / np.linalg.norm(ref.b)
        say(f"  {precond}: rc={rc} iters={st['krylov_iters']} relres={relres:.2e}")
        assert rc == 0 and relres <= 1e-8, (precond, relres, st)
        its[precond] = st["krylov_iters"]
        c.close()
    assert its["amg"] * (5 if dim == 2 else 2) <= its["jacobi"], its


@pytest.mark.parametrize("name", ["2d_dambreak", "3d_column"])
def test_picard_step_parity(ctx, name):
    mesh = small_meshes()[name]
    say(f"[picard] {name}")
    d, N = mesh.dim, mesh.n_nodes
    dt = 1e-3 if d == 2 else 5e-4
    q0 = mg.pspg_hydrostatic(mesh)
    # oracle: Picard with a tight Krylov stand-in
    pos_o = mesh.pos.reshape(-1).copy()
    q_o = q0.copy()
    c = make_case(ctx, mesh, krylov_rtol=1e-11)
    c.set_state(q0)
    for step in range(2):
        st_o, info = O.pspg_step(mesh, pos_o, q_o, dt=dt, solver_kind=2, rtol=1e-13)
        rc, st = c.pspg_step(dt)
        say(f"  step {step}: oracle iters={info['picard_iters']} gpu iters={st['picard_iters']} "
            f"krylov={st['krylov_iters']} res={st['residual']:.3e}")
        assert st_o == 0 and rc == 0
        assert abs(st["picard_iters"] - info["picard_iters"]) <= 1
        q_g = c.get_state()
        pos_g = c.get_positions().reshape(-1)
        assert block_rel_err(d, N, q_o, q_g) <= 1e-6
        np.testing.assert_allclose(pos_g, pos_o, rtol=0, atol=1e-9 * np.abs(pos_o).max())


@pytest.mark.parametrize("residual", ["Ax_f", "U", "U_P"])
def test_picard_residual_matches_oracle(ctx, residual):
    """minRes = 0 and maxIter = 1 fix the loop at exactly two Picard iterations (then
    iterCount > maxIter fails and restores): the residual after the second iteration is the
    same quantity in both, whichever residual kind is selected."""
    mesh = small_meshes()["2d_dambreak"]
    say(f"[picard residual] {residual}")
    dt = 1e-3
    q0 = mg.pspg_hydrostatic(mesh)
    pos_o = mesh.pos.reshape(-1).copy()
    q_o = q0.copy()
    st_o, info = O.pspg_step(mesh, pos_o, q_o, dt=dt, residual=residual, min_res=0.0, max_iter=1, solver_kind=2,
                             rtol=1e-13)
    c = make_case(ctx, mesh, residual=residual, min_res=0.0, max_iter=1, krylov_rtol=1e-12)
    c.set_state(q0)
    rc, st = c.pspg_step(dt)
    say(f"  oracle res={info['final_res']:.6e} iters={info['picard_iters']}  gpu res={st['residual']:.6e} "
        f"iters={st['picard_iters']}")
    assert st_o == 1 and rc == 1
    assert st["picard_iters"] == info["picard_iters"] == 2
    assert st["residual"] == pytest.approx(info["final_res"], rel=1e-5)


def test_picard_failure_restores(ctx):
    mesh = small_meshes()["2d_dambreak"]
    q0 = mg.pspg_hydrostatic(mesh)
    c = make_case(ctx, mesh, min_res=0.0, max_iter=0)
    c.set_state(q0)
    pos0 = c.get_positions().copy()
    rc, st = c.pspg_step(1e-3)
    assert rc == 1  # PFEM_NOT_CONVERGED: iterCount > maxIter
    np.testing.assert_array_equal(c.get_positions(), pos0)
    np.testing.assert_array_equal(c.get_state(), q0)


def test_invalid_arguments_raise(ctx):
    with pytest.raises(pfem_amd.PfemError):
        pfem_amd.Case(ctx, 4)
    with pytest.raises(pfem_amd.PfemError):
        pfem_amd.Case(ctx, 2, gamma=-0.07)  # a negative surface tension
    mesh = mg.dambreak_2d()
    bad = mg.PfemMesh(2, mesh.pos, mesh.conn.copy(), mesh.flags, mesh.h)
    bad.conn[0, 0] = mesh.n_nodes + 5
    c = pfem_amd.Case(ctx, 2)
    with pytest.raises(pfem_amd.PfemError):
        c.set_mesh(bad)
    # empty meshes, and meshes whose unknowns or pair-element incidence overflow 32-bit
    # indices, are rejected before anything is read from the arrays
    import ctypes

    lib = pfem_amd.lib()
    conn = np.zeros(3, np.int32)
    flags = np.zeros(1, np.uint8)
    pos = np.zeros(2)
    ptr = lambda a, t: a.ctypes.data_as(ctypes.POINTER(t))
    for n_nodes, n_elems in ((0, 1), (1, 0), (2**30, 1), (1, 2**28)):
        rc = lib.pfem_case_set_mesh(c._h, n_nodes, n_elems, ptr(conn, ctypes.c_int32), ptr(flags, ctypes.c_uint8),
                                    ptr(pos, ctypes.c_double))
        assert rc == pfem_amd._lib.PFEM_E_INVALID, (n_nodes, n_elems, rc)
    # the extractors need a mesh
    with pytest.raises(pfem_amd.PfemError):
        c.mass()
    with pytest.raises(pfem_amd.PfemError):
        c.minmax(0, "max")


def test_assembly_is_deterministic(ctx):
    """The gather assembly has no atomics: two cases with the same inputs give bitwise-identical
    A and b, and re-assembling gives the same bits again."""
    mesh = small_meshes()["3d_column"]
    q = rand_state(mesh, 11)
    rng = np.random.default_rng(12)
    dv = rng.normal(size=3 * mesh.n_nodes)
    out = []
    for _ in range(2):
        c = make_case(ctx, mesh)
        c.set_dirichlet(dv)
        c.assemble(5e-4, q_prev=q, q_iter=q)
        c.assemble(5e-4, q_prev=q, q_iter=q)
        out.append((c.csr()[2].copy(), c.rhs().copy()))
        c.close()
    assert np.array_equal(out[0][0], out[1][0])
    assert np.array_equal(out[0][1], out[1][1])


def derived_flags(mesh):
    """Node flags with FREE set on nodes that belong to no element, as the library derives them."""
    fl = mesh.flags.copy()
    used = np.zeros(mesh.n_nodes, bool)
    used[mesh.conn.reshape(-1)] = True
    fl[~used] |= mg.FREE
    return fl


@pytest.mark.parametrize("name", ["2d_dambreak", "3d_column"])
def test_extractors_match_oracle(ctx, name, tmp_path):
    """MassExtractor / MinMaxExtractor quantities on the device against the oracle, at the
    initial positions and after a Picard step has moved the nodes, and the CSV lines the
    extractors write for them."""
    from pfem_amd.extractors import MassExtractor, MinMaxExtractor

    mesh = small_meshes()[name]
    say(f"[extract] {name}")
    d = mesh.dim
    c = pfem_amd.MomContEqIncompNewton(ctx, d, rho=1000.0, mu=1e-3, residual="Ax_f", min_res=1e-6, max_iter=10)
    c.set_mesh(mesh)
    c.set_state(mg.pspg_hydrostatic(mesh))
    fl = derived_flags(mesh)
    dt = 1e-3 if d == 2 else 5e-4
    ex_mass = MassExtractor(c, str(tmp_path / "mass.csv"), dt)
    ex_top = MinMaxExtractor(c, str(tmp_path / "top.csv"), dt, d - 1, "max")
    expect_mass, expect_top = [], []
    for step in range(2):
        pos = c.get_positions().reshape(d, -1)
        m_ref = O.mass(mesh, pos)
        m_gpu = c.mass()
        say(f"  step {step}: mass {m_gpu:.15g} vs {m_ref:.15g}")
        assert m_gpu == pytest.approx(m_ref, rel=1e-12)
        for k in range(d):
            for kind in ("min", "max"):
                assert c.minmax(k, kind) == O.minmax(mesh, pos, k, kind, flags=fl), (k, kind)
        t = step * dt
        assert ex_mass.update(t) and ex_top.update(t)
        expect_mass.append((t, m_ref))
        expect_top.append("%f,%f" % (t, O.minmax(mesh, pos, d - 1, "max", flags=fl)))
        if step == 0:
            assert c.solve(dt)
    ex_mass.close()
    ex_top.close()
    lines = (tmp_path / "mass.csv").read_text().splitlines()
    assert len(lines) == len(expect_mass)
    for line, (t, m) in zip(lines, expect_mass):
        ts, ms = line.split(",")
        assert ts == "%f" % t and float(ms) == pytest.approx(m, rel=1e-12, abs=5e-7)
    assert (tmp_path / "top.csv").read_text().splitlines() == expect_top
    with pytest.raises(pfem_amd.PfemError):
        c.minmax(d, "max")
    c.close()


@pytest.mark.parametrize("name", ["2d_dambreak", "2d_column", "3d_column"])
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_ell_copy_matches_node_block_values(ctx, name, precision):
    """The sliced ELL copy the BiCGSTAB SpMVs read, written by the assembly gathers and the
    BC kernel beside the CSR-order values, gives the same A x (free nodes, boundary rows,
    Dirichlet column elimination and the slices' padding ranks included)."""
    mesh = small_meshes()[name]
    d, N = mesh.dim, mesh.n_nodes
    rng = np.random.default_rng(8)
    q = rand_state(mesh, 3)
    dv = rng.normal(scale=0.1, size=d * N)
    say(f"[ell copy] {name} {precision} N={N}")
    c = pfem_amd.Case(ctx, d, kind="PSPG", precision=precision, precond="amg" if precision == "mixed" else "auto")
    c.set_mesh(mesh)
    c.set_dirichlet(dv)
    for it in range(2):  # a second assembly overwrites every entry of the first
        c.assemble(1e-3 if d == 2 else 5e-4, q_prev=q, q_iter=q * (1.0 + 0.1 * it))
        x = rng.normal(size=(d + 1) * N)
        y_csr, y_ell = c.spmv(x), c.spmv_ell(x)
        scale = np.abs(y_csr).max()
        assert np.abs(y_ell - y_csr).max() <= 1e-13 * scale
    c.close()


@pytest.mark.parametrize("name", ["2d_dambreak", "3d_column"])
def test_bingham_assembly_and_step_parity(ctx, name):
    """Bingham fluid (MomContEquation.inl:100-123): the K factor of every element is
    mu + mu_eq of its strain rate in the current iterate; assembly and a Picard step against
    the oracle."""
    mesh = small_meshes()[name]
    d, N = mesh.dim, mesh.n_nodes
    dt = 1e-3 if d == 2 else 5e-4
    tau0, mreg = 5.0, 200.0
    qprev, qiter = rand_state(mesh, 21), rand_state(mesh, 22)
    say(f"[bingham] {name} N={N}")
    with O.material("Bingham", tau0=tau0, mReg=mreg):
        ref = O.pspg_assemble(mesh, qprev, qiter, dt=dt)
    c = pfem_amd.Case(ctx, d, kind="PSPG", problem="Bingham", tau0=tau0, mReg=mreg, krylov_rtol=1e-11)
    c.set_mesh(mesh)
    c.assemble(dt, q_prev=qprev, q_iter=qiter)
    rp, ci, va = c.csr()
    np.testing.assert_array_equal(rp, ref.row_ptr)
    np.testing.assert_array_equal(ci, ref.col_idx)
    assert row_rel_err(rp, ref.vals, va) <= 1e-12
    assert block_rel_err(d, N, ref.b, c.rhs()) <= 1e-12
    # the Newtonian system differs (the yield stress matters on this field)
    newt = O.pspg_assemble(mesh, qprev, qiter, dt=dt)
    assert row_rel_err(rp, newt.vals, va) > 1e-6
    c.close()
    # one Picard step from the hydrostatic state
    q0 = mg.pspg_hydrostatic(mesh)
    # the viscosity follows the iterate, so the loop needs more passes than a Newtonian one
    # (about 21 on the 2D case at minRes 1e-6)
    eq = pfem_amd.MomContEqIncompNewton(ctx, d, problem="Bingham", tau0=tau0, mReg=mreg, krylov_rtol=1e-11,
                                        max_iter=30)
    eq.set_mesh(mesh)
    eq.set_state(q0)
    assert eq.solve(dt), eq.last_stats
    pos = mesh.pos.reshape(-1).copy()
    q = q0.copy()
    with O.material("Bingham", tau0=tau0, mReg=mreg):
        st, info = O.pspg_step(mesh, pos, q, dt=dt, solver_kind=2, rtol=1e-13, max_iter=30)
    assert st == 0
    assert abs(eq.last_stats["picard_iters"] - info["picard_iters"]) <= 1
    assert block_rel_err(d, N, q, eq.get_state()) <= 1e-6
    eq.close()
