"""Heat equation (HeatEqIncompNewton without phase change, the Conduction problem) on the GPU
against the oracle (oracle/pfem_oracle.cpp heat section): the assembled system with its facet
terms and Dirichlet elimination (pattern exact, values and rhs to 1e-12), and time steps of
the conjugate-gradient solve (temperatures to 1e-9), in 2D and 3D, with prescribed flux,
convection and radiation facets."""
import itertools
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


def block_case(dim, n, h=0.1, seed=0):
    rng = np.random.default_rng(seed)
    P = np.array(list(itertools.product(*[range(n)] * dim)), dtype=float) * h
    inner = ((P > 0) & (P < (n - 1) * h - 1e-12)).all(axis=1)
    P[inner] += rng.uniform(-0.15 * h, 0.15 * h, size=(inner.sum(), dim))
    P = np.vstack([P, np.full((1, dim), 10.0)])          # a node in no element (free)
    m = pfem_amd.Mesher(dim, hchar=h, alpha=1.3)
    m.set_nodes(P.T.copy(), np.zeros(len(P), np.uint8))
    m.triangulate()
    pos, _, fl, _ = m.nodes()
    fac, out, _ = m.facets()
    fl = fl.copy()
    # Dirichlet on the x = 0 side; convection + radiation on the others, a flux on the top
    bnd = np.unique(fac)
    left = bnd[pos[0, bnd] < 1e-12]
    fl[left] |= mg.DIRICHLET
    mesh = mg.PfemMesh(dim, pos, m.elements(), fl, h)
    fpos = pos[:, fac]                                     # dim x F x dim
    on_left = (np.abs(fpos[0]) < 1e-12).all(axis=1)
    top = (np.abs(fpos[dim - 1] - (n - 1) * h) < 1e-12).all(axis=1)
    kinds = np.where(on_left, 0, O.HEAT_QH | O.HEAT_QR).astype(np.uint8)
    kinds[top] |= O.HEAT_Q
    return mesh, fac, out, kinds, left


@pytest.mark.parametrize("dim,n", [(2, 9), (3, 5)])
def test_heat_assembly_and_steps_match_oracle(ctx, dim, n):
    mesh, fac, out, kinds, left = block_case(dim, n)
    N = mesh.n_nodes
    say(f"[heat] {dim}D N={N} E={mesh.n_elems} F={len(fac)}")
    rng = np.random.default_rng(3)
    T = 300.0 + rng.normal(scale=10.0, size=N)
    Tbc = np.zeros(N)
    Tbc[left] = 400.0 + 20.0 * mesh.pos[1, left]
    mat = dict(rho=900.0, cv=2000.0, k=5.0, h=30.0, Tinf=290.0, epsRad=0.7)
    case = pfem_amd.HeatEqIncompNewton(ctx, dim, krylov_rtol=1e-13, **mat)
    case.set_mesh(mesh)
    case.set_state(T)
    case.set_facets(fac, out, kinds)
    xg = case.facet_points()
    # the Gauss points: on the facet, the mean of a facet's points is its centroid
    np.testing.assert_allclose(xg.mean(axis=1), mesh.pos[:, fac].transpose(1, 2, 0).mean(axis=1), rtol=0,
                               atol=1e-13)
    q = np.zeros((len(fac), 3, dim))
    q[..., dim - 1] = -50.0 + 5.0 * xg[..., 0]            # "TopQ"(pos): heat entering through the top
    case.set_flux(q)
    case.set_dirichlet_T(Tbc)
    hc = O.HeatCase(facets=fac, facet_out=out, facet_kind=kinds, qgp=q, **mat)
    dt = 5.0
    case.assemble(dt)
    rp, ci, v, b = case.system()
    S = O.heat_assemble(mesh, T, Tbc=Tbc, hc=hc, dt=dt)
    np.testing.assert_array_equal(rp, S.row_ptr)
    np.testing.assert_array_equal(ci, S.col_idx)
    row = np.repeat(np.arange(N), np.diff(rp))
    rmax = np.zeros(N)
    np.maximum.at(rmax, row, np.abs(S.vals))
    assert (np.abs(v - S.vals) / np.maximum(rmax[row], 1e-300)).max() <= 1e-12
    assert np.abs(b - S.b).max() <= 1e-12 * np.abs(S.b).max()
    Tg, To = T.copy(), T.copy()
    for step in range(3):
        assert case.solve(dt), case.last_stats
        st, it, rr = O.heat_step(mesh, To, Tbc=Tbc, hc=hc, dt=dt, rtol=1e-13)
        assert st == 0
        Tg = case.get_state()
        say(f"  step {step}: gpu CG {case.last_stats['krylov_iters']} its (relres "
            f"{case.last_stats['max_relres']:.1e}), oracle {it} its; max |dT| {np.abs(Tg - To).max():.2e}")
        assert np.abs(Tg - To).max() <= 1e-9 * np.abs(To).max()
        assert Tg[-1] == T[-1]                              # the free node keeps its temperature
        # Dirichlet nodes take their values (to the solve's tolerance: their unit rows are solved too)
        np.testing.assert_allclose(Tg[left], Tbc[left], rtol=1e-12)
    case.close()


def test_heat_case_is_refused_by_the_batch_and_rejects_bad_input(ctx):
    mesh, fac, out, kinds, _ = block_case(2, 5)
    case = pfem_amd.HeatEqIncompNewton(ctx, 2)
    case.set_mesh(mesh)
    with pytest.raises(pfem_amd.PfemError) as ei:
        pfem_amd.batch_step(ctx, [case], [1.0])
    assert ei.value.code == -4
    bad = fac.copy()
    bad[0, 0] = mesh.n_nodes
    with pytest.raises(pfem_amd.PfemError):
        case.set_facets(bad, out, kinds)
    with pytest.raises(pfem_amd.PfemError):
        pfem_amd.HeatEqIncompNewton(ctx, 2, cv=0.0)
    case.close()
