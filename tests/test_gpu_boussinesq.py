"""The Boussinesq problem (PSPG momentum-continuity with the temperature-dependent body force,
and the heat equation, IncompNewton/Solver.cpp:249-283) on the GPU against the oracle
(oracle.boussinesq_step): a heated cavity whose bottom and top hold Dirichlet temperatures and
whose side walls lose heat by convection, both equation orders, velocities and pressure (up to
its constant: a closed cavity) to 1e-7 of their scale, temperatures to 1e-9, positions to 1e-9
of the domain; and the refusals."""
import sys

import numpy as np
import pytest

import pfem_amd
from oracle import oracle as O
from pfem_amd import _lib
from pfem_amd import meshgen as mg
from test_oracle_heat_pins import block_mesh

pytestmark = pytest.mark.gpu

BOUND, FIXED, DIRICHLET, DIRICHLET_T = 1, 8, 16, 32


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def heated_cavity(dim):
    mesh, fac, out = block_mesh(dim, n=8 if dim == 2 else 5, seed=11)
    P = mesh.pos.T
    hi = P.max(axis=0)
    wall = ((P < 1e-12) | (P > hi - 1e-12)).any(axis=1)
    bottom = P[:, dim - 1] < 1e-12
    top = P[:, dim - 1] > hi[dim - 1] - 1e-12
    flags = np.where(wall, BOUND | FIXED | DIRICHLET, 0).astype(np.uint8)
    flags[bottom | top] |= DIRICHLET_T
    mesh.flags = flags
    # convection on the facets of the side walls (every node on a side wall)
    side = ((P[:, : dim - 1] < 1e-12) | (P[:, : dim - 1] > hi[: dim - 1] - 1e-12)).any(axis=1)
    kinds = np.where(side[fac].all(axis=1), O.HEAT_QH, 0).astype(np.uint8)
    Tbc = np.where(bottom, 360.0, np.where(top, 300.0, 0.0))
    rng = np.random.default_rng(12)
    T0 = np.where(bottom, 360.0, np.where(top, 300.0, 330.0 + rng.normal(scale=5.0, size=len(P))))
    return mesh, fac, out, kinds, Tbc, T0


@pytest.mark.parametrize("dim,heat_first", [(2, True), (2, False), (3, True)])
def test_boussinesq_steps_match_oracle(ctx, dim, heat_first):
    mesh, fac, out, kinds, Tbc, T0 = heated_cavity(dim)
    N = mesh.n_nodes
    say(f"[boussinesq] {dim}D heat_first={heat_first} N={N} E={mesh.n_elems} F={len(fac)}")
    mat = dict(rho=1000.0, mu=1e-2, cv=4000.0, k=50.0, alpha=2e-3, Tr=330.0, h=20.0, Tinf=290.0)
    g = [0.0, -9.81] if dim == 2 else [0.0, 0.0, -9.81]
    dt = 5e-3
    q = np.concatenate([np.zeros(dim * N), 1000.0 * (mesh.pos.T @ np.array(g)), T0])
    case = pfem_amd.BoussinesqIncompNewton(ctx, dim, solve_heat_first=heat_first, body_force=g, residual="U",
                                           min_res=1e-9, max_iter=30, krylov_rtol=1e-12, **mat)
    case.set_mesh(mesh)
    case.set_state(q)
    case.set_facets(fac, out, kinds)
    case.set_dirichlet(np.zeros((dim, N)))
    case.set_dirichlet_T(Tbc)
    hc = O.HeatCase(rho=mat["rho"], cv=mat["cv"], k=mat["k"], h=mat["h"], Tinf=mat["Tinf"], facets=fac,
                    facet_out=out, facet_kind=kinds)
    q_o, pos_o = q.copy(), mesh.pos.copy()
    nm = (dim + 1) * N
    for step in range(3):
        st_o, which = O.boussinesq_step(mesh, pos_o, q_o, hc, mat["alpha"], mat["Tr"], Tbc=Tbc, rho=mat["rho"],
                                        mu=mat["mu"], g=g, dt=dt, solve_heat_first=heat_first, residual="U",
                                        min_res=1e-9, max_iter=30, rtol=1e-13)
        ok = case.solve(dt)
        say(f"  step {step}: oracle {st_o} {which} gpu {ok} {case.failed} "
            f"picard={case.last_stats['momentum']['picard_iters']}")
        assert st_o == 0 and ok
        qg = case.get_state()
        vscale = max(np.abs(q_o[: dim * N]).max(), 1e-12)
        assert np.abs(qg[: dim * N] - q_o[: dim * N]).max() < 1e-7 * vscale
        # every wall holds a Dirichlet velocity: the pressure is fixed up to a constant, which
        # each solver (BiCGSTAB here, BiCGSTAB + ILU(0) in the oracle) picks for itself
        assert np.ptp(qg[dim * N: nm] - q_o[dim * N: nm]) < 1e-7 * np.abs(q_o[dim * N: nm]).max()
        assert np.abs(qg[nm:] - q_o[nm:]).max() < 1e-9 * np.abs(q_o[nm:]).max()
        assert np.abs(case.get_positions() - pos_o).max() < 1e-9
    # the buoyancy has set the fluid in motion
    assert np.abs(q_o[: dim * N]).max() > 0
    case.close()


def test_boussinesq_refusals(ctx):
    with pytest.raises(pfem_amd.PfemError) as e:
        pfem_amd.BoussinesqIncompNewton(ctx, 2, DgammaDT=-1e-3)
    assert e.value.code == _lib.PFEM_E_INVALID
    with pytest.raises(pfem_amd.PfemError) as e:
        pfem_amd.BoussinesqIncompNewton(ctx, 2, problem="Bingham", tau0=1.0, mReg=10.0)
    assert e.value.code == _lib.PFEM_E_INVALID
    mesh, fac, out, kinds, Tbc, T0 = heated_cavity(2)
    case = pfem_amd.BoussinesqIncompNewton(ctx, 2)
    case.set_mesh(mesh)
    q = np.arange(4 * mesh.n_nodes, dtype=float)
    case.set_state(q)
    assert np.array_equal(case.get_state(), q)
    with pytest.raises(pfem_amd.PfemError) as e:
        pfem_amd.batch_step(ctx, [case], [1e-3])
    assert e.value.code == _lib.PFEM_E_UNSUPPORTED
    case.close()
