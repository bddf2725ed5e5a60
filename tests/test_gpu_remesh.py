"""Time steps with remeshing between them (SolverIncompNewton::m_solveIncompNewtonNoT,
IncompNewton/Solver.cpp:236-246): the GPU step followed by pfem_case_remesh (positions and node
states to the host mesher, the new mesh and states back) against the oracle's step followed by
the same mesher on the oracle's positions and states.  The mesher itself is checked against the
remeshing oracle in tests/test_mesher.py; here the hand-over is: node counts, elements and flags
equal after every remesh, states and positions to the step tolerance.  The mesh parameters are
the 2D Koshizuka dam break's (examples/2D/damBreakKoshizuka/damBreakKoshizukaIncomp.lua:6-15)."""
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


def mesher_for(mesh, **kw):
    prm = dict(hchar=mesh.h, alpha=1.2, omega=0.7, gamma=0.2, addOnFS=False, deleteFlyingNodes=True,
               laplacianSmoothingBoundaries=True)
    prm.update(kw)
    return pfem_amd.Mesher(mesh.dim, **prm)


@pytest.mark.parametrize("dim", [2, 3])
def test_steps_with_remesh_match_oracle(ctx, dim):
    """Every step starts from the remeshed mesh: the GPU step against the oracle's step on that
    mesh and state, then pfem_case_remesh against a second mesher fed the positions and states
    the case hands over (a Delaunay triangulation can change with the last bits of the
    positions, so the two remeshes are compared on the same input, bit for bit)."""
    mesh = mg.dambreak_2d(h=0.1) if dim == 2 else mg.dambreak_3d(h=0.1, col=(0.5, 0.5, 1.0))
    d, N = mesh.dim, mesh.n_nodes
    dt = 1e-3 if d == 2 else 5e-4
    q0 = mg.pspg_hydrostatic(mesh)
    say(f"[remesh steps] {d}D N={N}")
    # omega and gamma tighter than the example's, so that nodes are added and removed every step.
    # (In 3D an omega below h^3/6, the Kuhn tet's volume, splits the elements between the
    # un-jittered wall lattices too: their centroids are then nearly co-planar and the Delaunay
    # triangulation holds slivers of detJ ~ 1e-17 h^3 that no step survives, as it would for the
    # reference's CGAL triangulation of the same nodes.)
    kw = dict(omega=0.45, gamma=0.6) if dim == 2 else dict(omega=0.2, gamma=0.5)
    m_gpu, m_chk = mesher_for(mesh, **kw), mesher_for(mesh, **kw)
    for m in (m_gpu, m_chk):
        m.set_nodes(mesh.pos, mesh.flags, q0.reshape(d + 1, N))
        m.triangulate()
    case = pfem_amd.MomContEqIncompNewton(ctx, d, krylov_rtol=1e-12)
    case.set_mesh_from(m_gpu)
    changed = 0
    for step in range(3):
        om = m_gpu.mesh(h=mesh.h)
        _, st, _, _ = m_gpu.nodes()
        pos_o, q_o = om.pos.reshape(-1).copy(), st.reshape(-1).copy()
        ok_o, info = O.pspg_step(om, pos_o, q_o, dt=dt, solver_kind=2, rtol=1e-13)
        assert ok_o == 0
        assert case.solve(dt), case.last_stats
        q_g, pos_g = case.get_state(), case.get_positions()
        Nn = om.n_nodes
        for s in range(d + 1):
            ref = q_o[s * Nn:(s + 1) * Nn]
            assert np.abs(q_g[s * Nn:(s + 1) * Nn] - ref).max() <= 1e-6 * max(np.abs(ref).max(), 1e-30), (step, s)
        np.testing.assert_allclose(pos_g.reshape(-1), pos_o, rtol=0, atol=1e-10)
        # remesh: the case hands its nodes to m_gpu; m_chk gets the same numbers from the host
        case.remesh(m_gpu)
        m_chk.set_positions_states(pos_g, q_g.reshape(d + 1, -1))
        m_chk.remesh()
        sg = m_gpu.stats()
        changed += sg["added"] + sg["removed"] + sg["out_of_box"] + sg["flying_deleted"]
        say(f"  step {step}: picard={info['picard_iters']:.0f} gpu={case.last_stats['picard_iters']} "
            f"-> N={m_gpu.sizes()[0]} E={m_gpu.sizes()[1]} added={sg['added']:.0f} removed={sg['removed']:.0f}")
        assert m_gpu.sizes() == m_chk.sizes()
        np.testing.assert_array_equal(m_gpu.elements(), m_chk.elements())
        a, b = m_gpu.nodes(), m_chk.nodes()
        for x, y in zip(a, b):
            np.testing.assert_array_equal(x, y)
        # the case holds the remeshed mesh and states
        assert case.N == m_gpu.sizes()[0]
        np.testing.assert_array_equal(case.get_positions(), a[0])
        np.testing.assert_array_equal(case.get_state().reshape(d + 1, -1), a[1])
    assert changed > 0
    case.close()


def test_solver_loop_steps_and_adapts_dt(ctx):
    """SolverIncompNewton: step then remesh; dt grows by coeffDTincrease up to maxDT."""
    mesh = mg.dambreak_2d(h=0.1)
    m = mesher_for(mesh)
    m.set_nodes(mesh.pos, mesh.flags, mg.pspg_hydrostatic(mesh).reshape(3, -1))
    m.triangulate()
    case = pfem_amd.MomContEqIncompNewton(ctx, 2)
    case.set_mesh_from(m)
    solver = pfem_amd.SolverIncompNewton(case, m, initial_dt=1e-3, max_dt=2e-3, coeff_dt_increase=1.5)
    dts = []
    for _ in range(3):
        dts.append(solver.dt)
        assert solver.solveOneTimeStep()
        solver.computeNextDT()
    assert dts == [1e-3, 1.5e-3, 2e-3]
    assert solver.time == pytest.approx(4.5e-3, abs=1e-15)
    assert case.N == m.sizes()[0]
    case.close()


def test_remesh_refuses_foreign_mesher(ctx):
    mesh = mg.dambreak_2d(h=0.1)
    case = pfem_amd.MomContEqIncompNewton(ctx, 2)
    case.set_mesh(mesh)
    case.set_state(mg.pspg_hydrostatic(mesh))
    m = mesher_for(mesh)
    with pytest.raises(pfem_amd.PfemError):
        case.remesh(m)   # the mesher holds no nodes of this case
    m.set_nodes(mesh.pos, mesh.flags, np.zeros((4, mesh.n_nodes)))
    m.triangulate()
    with pytest.raises(pfem_amd.PfemError):
        case.remesh(m)   # 4 states per node is not a 2D PSPG case's 3
    case.close()


def test_multigrid_on_a_3d_delaunay_mesh(ctx):
    """A Delaunay tetrahedrisation of a node cloud holds slivers; the multigrid then smooths
    with damping 0.5 (at 0.9 BiCGSTAB diverged on such meshes) and the step matches the
    oracle's."""
    mesh = mg.dambreak_3d(h=0.033, col=(0.5, 0.5, 1.0))
    q0 = mg.pspg_hydrostatic(mesh)
    m = mesher_for(mesh, laplacianSmoothingBoundaries=False)
    m.set_nodes(mesh.pos, mesh.flags, q0.reshape(4, -1))
    m.triangulate()
    case = pfem_amd.MomContEqIncompNewton(ctx, 3, precond="amg", krylov_rtol=1e-11, krylov_max_iter=3000)
    case.set_mesh_from(m)
    st = case.setup_times()
    say(f"[delaunay amg] N={case.N} E={case.E} min quality {st['min_quality']:.2e} omega {st['amg_omega']}")
    assert st["min_quality"] < 0.1 and st["amg_omega"] == 0.5
    assert case.solve(5e-4), case.last_stats
    assert case.last_stats["krylov_iters"] < 600
    om = m.mesh(h=mesh.h)
    _, s0, _, _ = m.nodes()
    pos, q = om.pos.reshape(-1).copy(), s0.reshape(-1).copy()
    ok, info = O.pspg_step(om, pos, q, dt=5e-4, solver_kind=2, rtol=1e-13)
    assert ok == 0 and abs(info["picard_iters"] - case.last_stats["picard_iters"]) <= 1
    g = case.get_state()
    N = case.N
    for s in range(4):
        ref = q[s * N:(s + 1) * N]
        assert np.abs(g[s * N:(s + 1) * N] - ref).max() <= 1e-6 * np.abs(ref).max()
    # a lattice-split mesh keeps the faster damping
    case.set_mesh(mesh)
    assert case.setup_times()["amg_omega"] == 0.9
    case.close()
