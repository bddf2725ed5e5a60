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
    mesh = mg.dambreak_2d(h=0.1) if dim == 2 else mg.dambreak_3d(h=0.1, col=(0.5, 0.5, 1.0))
    d, N = mesh.dim, mesh.n_nodes
    dt = 1e-3 if d == 2 else 5e-4
    q0 = mg.pspg_hydrostatic(mesh)
    say(f"[remesh steps] {d}D N={N}")
    # omega and gamma tighter than the example's, so that nodes are added and removed every step
    kw = dict(omega=0.45 if dim == 2 else 0.12, gamma=0.6)
    # the GPU path
    m_gpu = mesher_for(mesh, **kw)
    m_gpu.set_nodes(mesh.pos, mesh.flags, q0.reshape(d + 1, N))
    m_gpu.triangulate()
    case = pfem_amd.MomContEqIncompNewton(ctx, d, krylov_rtol=1e-12)
    case.set_mesh_from(m_gpu)
    solver = pfem_amd.SolverIncompNewton(case, m_gpu, initial_dt=dt, max_dt=dt)
    # the oracle path through its own mesher instance
    m_orc = mesher_for(mesh, **kw)
    m_orc.set_nodes(mesh.pos, mesh.flags, q0.reshape(d + 1, N))
    m_orc.triangulate()
    changed = 0
    for step in range(3):
        om = m_orc.mesh(h=mesh.h)
        pos = om.pos.reshape(-1).copy()
        _, st, _, _ = m_orc.nodes()
        q = st.reshape(-1).copy()
        ok_o, info = O.pspg_step(om, pos, q, dt=dt, solver_kind=2, rtol=1e-13)
        assert ok_o == 0
        m_orc.set_positions_states(pos.reshape(d, -1), q.reshape(d + 1, -1))
        m_orc.remesh()
        assert solver.solveOneTimeStep(), case.last_stats if hasattr(case, "last_stats") else None
        solver.computeNextDT()
        so, sg = m_orc.stats(), m_gpu.stats()
        changed += so["added"] + so["removed"] + so["out_of_box"] + so["flying_deleted"]
        say(f"  step {step}: N={m_gpu.sizes()[0]} E={m_gpu.sizes()[1]} added={sg['added']:.0f} "
            f"removed={sg['removed']:.0f} flying={sg['flying_deleted']:.0f} picard={info['picard_iters']:.0f}")
        assert m_gpu.sizes() == m_orc.sizes()
        np.testing.assert_array_equal(m_gpu.elements(), m_orc.elements())
        pg, sgs, fg, _ = m_gpu.nodes()
        po, sos, fo, _ = m_orc.nodes()
        np.testing.assert_array_equal(fg, fo)
        np.testing.assert_allclose(pg, po, rtol=0, atol=1e-10)
        for s in range(d + 1):
            ref = sos[s]
            assert np.abs(sgs[s] - ref).max() <= 1e-6 * max(np.abs(ref).max(), 1e-30), (step, s)
        # the case holds the remeshed mesh and states
        assert case.N == m_gpu.sizes()[0]
        np.testing.assert_array_equal(case.get_positions(), pg)
        np.testing.assert_array_equal(case.get_state().reshape(d + 1, -1), sgs)
    assert changed > 0
    assert abs(solver.time - 3 * dt) < 1e-15
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
