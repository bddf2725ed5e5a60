"""Surface tension (the FST facet term) on the GPU against the oracle: the PSPG right-hand side of
jittered polygon / icosahedron drops (1e-12), Picard steps of a 2D drop at rest triangulated by
the mesher with its facets handed over by pfem_case_set_mesh_from (pressure to 1e-6, velocities
to 1e-9 of the capillary velocity, +-1 Picard iteration), a WC step of the drops (accelerations
to 1e-10), and the refusals."""
import sys

import numpy as np
import pytest

import pfem_amd
from oracle import oracle as O
from pfem_amd import _lib
from test_oracle_fst_pins import icosa_drop, polygon_drop

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def drop(dim):
    return polygon_drop(11, 0.2, jitter=1.0, seed=21) if dim == 2 else icosa_drop(0.2, jitter=1.0, seed=22)


@pytest.mark.parametrize("dim", [2, 3])
def test_pspg_fst_rhs_matches_oracle(ctx, dim):
    mesh, fac, elem = drop(dim)
    N = mesh.n_nodes
    say(f"[fst] PSPG rhs {dim}D N={N} F={len(fac)}")
    rng = np.random.default_rng(5)
    q = np.concatenate([0.01 * rng.standard_normal(dim * N), 50.0 * rng.standard_normal(N)])
    g, dt, gamma = [0.0] * dim, 1e-3, 0.072
    c = pfem_amd.Case(ctx, dim, kind="PSPG", gamma=gamma, body_force=g)
    c.set_mesh(mesh)
    c.set_surface_facets(fac, elem)
    c.assemble(dt, q_prev=q, q_iter=q)
    b = c.rhs()
    with O.surface_tension(fac, elem, gamma):
        ref = O.pspg_assemble(mesh, q, q, g=g, dt=dt)
    base = O.pspg_assemble(mesh, q, q, g=g, dt=dt)
    assert np.abs(ref.b - base.b).max() > 1e-3 * gamma  # the term is there
    for s in range(dim + 1):
        blk = slice(s * N, (s + 1) * N)
        assert np.abs(b[blk] - ref.b[blk]).max() <= 1e-12 * max(np.abs(ref.b[blk]).max(), 1e-300)
    c.close()


@pytest.mark.parametrize("dim", [2, 3])
def test_boussinesq_fst_rhs_matches_oracle(ctx, dim):
    """The Boussinesq factor gamma + DgammaDT (T - Tr) at the facet Gauss points."""
    mesh, fac, elem = drop(dim)
    N = mesh.n_nodes
    rng = np.random.default_rng(8)
    q = np.concatenate([0.01 * rng.standard_normal(dim * N), 50.0 * rng.standard_normal(N)])
    T = 300.0 + 40.0 * rng.random(N)
    g, dt, gamma, dgdT, alpha, Tr = [0.0] * dim, 1e-3, 0.072, 2e-4, 1e-3, 310.0
    c = pfem_amd.BoussinesqIncompNewton(ctx, dim, gamma=gamma, DgammaDT=dgdT, alpha=alpha, Tr=Tr, body_force=g)
    c.set_mesh(mesh)
    c.set_state(np.concatenate([q, T]))
    c.set_surface_facets(fac, elem)
    c.assemble(dt, q_prev=q, q_iter=q)
    b = c.rhs()
    with O.material("Boussinesq", alpha=alpha, Tr=Tr, T=T), O.surface_tension(fac, elem, gamma, dgdT):
        ref = O.pspg_assemble(mesh, q, q, g=g, dt=dt)
    with O.material("Boussinesq", alpha=alpha, Tr=Tr, T=T), O.surface_tension(fac, elem, gamma, 0.0):
        flat = O.pspg_assemble(mesh, q, q, g=g, dt=dt)
    assert np.abs(ref.b - flat.b).max() > 1e-3 * gamma  # DgammaDT changes the rhs
    for s in range(dim + 1):
        blk = slice(s * N, (s + 1) * N)
        assert np.abs(b[blk] - ref.b[blk]).max() <= 1e-12 * max(np.abs(ref.b[blk]).max(), 1e-300)
    c.close()


def disk_drop_mesher(R=0.05, h=0.01):
    pts = [[0.0, 0.0]]
    for k in range(1, int(round(R / h)) + 1):
        r = k * h
        n = max(6, int(round(2 * np.pi * r / h)))
        th = 2 * np.pi * np.arange(n) / n + 0.3 * k
        pts += list(np.stack([r * np.cos(th), r * np.sin(th)], axis=1))
    P = np.array(pts).T.copy()
    m = pfem_amd.Mesher(2, hchar=h, alpha=1.3)
    m.set_nodes(P, np.zeros(P.shape[1], np.uint8), states=np.zeros((3, P.shape[1])))
    m.triangulate()
    return m


def test_pspg_fst_drop_steps_match_oracle(ctx):
    m = disk_drop_mesher()
    mesh = m.mesh(h=0.01)
    fac, _, elem = m.facets()
    N = mesh.n_nodes
    say(f"[fst] 2D drop N={N} E={mesh.n_elems} F={len(fac)} onfs={(mesh.flags & 4 > 0).sum()}")
    gamma, dt = 0.072, 1e-3
    c = pfem_amd.Case(ctx, 2, kind="PSPG", gamma=gamma, body_force=[0.0, 0.0], residual="Ax_f", krylov_rtol=1e-11)
    c.set_mesh_from(m)  # hands over the mesher's facets
    q0 = np.zeros(3 * N)
    c.set_state(q0)
    q_o, pos_o = q0.copy(), mesh.pos.copy()
    for step in range(2):
        with O.surface_tension(fac, elem, gamma):
            st_o, info = O.pspg_step(mesh, pos_o, q_o, g=[0.0, 0.0], dt=dt, solver_kind=2, rtol=1e-13)
        rc, st = c.pspg_step(dt)
        say(f"  step {step}: oracle {st_o} iters {info['picard_iters']} gpu {rc} iters {st['picard_iters']}")
        assert st_o == 0 and rc == 0
        assert abs(st["picard_iters"] - info["picard_iters"]) <= 1
        qg = c.get_state()
        # the drop starts at rest and the surface tension is balanced by the pressure jump, so
        # the velocities stay at rounding level: they are compared against the capillary
        # velocity scale sqrt(gamma / (rho R)) ~ 0.04 m/s, the pressure against itself
        vs = np.sqrt(gamma / (1000.0 * 0.05))
        assert np.abs(qg[: 2 * N] - q_o[: 2 * N]).max() <= 1e-9 * vs
        assert np.abs(qg[2 * N:] - q_o[2 * N:]).max() <= 1e-6 * np.abs(q_o[2 * N:]).max()
        assert np.abs(c.get_positions() - pos_o).max() <= 1e-9 * np.abs(pos_o).max()
    # the surface tension has set the drop's pressure: about gamma / R inside
    p = c.get_state()[2 * N:]
    assert 0.3 * gamma / 0.05 < p[0] < 3 * gamma / 0.05
    c.close()


@pytest.mark.parametrize("dim", [2, 3])
def test_wc_fst_step_matches_oracle(ctx, dim):
    mesh, fac, elem = drop(dim)
    N = mesh.n_nodes
    say(f"[fst] WC {dim}D N={N}")
    gamma, dt, rho_star = 0.06, 1e-6, 1000.0
    v = np.zeros(dim * N)
    p = np.zeros(N)
    rho = np.full(N, rho_star)
    a = np.zeros(dim * N)
    c = pfem_amd.Case(ctx, dim, kind="WC", gamma=gamma, body_force=[0.0] * dim, mu=1e-3, rho_star=rho_star)
    c.set_mesh(mesh)
    c.set_surface_facets(fac, elem)
    c.set_state(np.concatenate([v, p, rho, a]))
    c.wc_step(dt)
    qg = c.get_state()
    pos = mesh.pos.copy()
    with O.surface_tension(fac, elem, gamma):
        O.wc_step(mesh, pos, v, p, rho, a, dt, mu=1e-3, g=[0.0] * dim, rho_star=rho_star)
    ag = qg[(dim + 2) * N:]
    assert np.abs(a).max() > 0
    assert np.abs(ag - a).max() <= 1e-10 * np.abs(a).max()
    assert np.abs(qg[: dim * N] - v).max() <= 1e-10 * np.abs(v).max()
    c.close()


def test_fst_refusals(ctx):
    with pytest.raises(pfem_amd.PfemError) as e:
        pfem_amd.MomContEqFracStep(ctx, 2, gamma=0.07)
    assert e.value.code == _lib.PFEM_E_UNSUPPORTED
    mesh, fac, elem = drop(2)
    c = pfem_amd.Case(ctx, 2, kind="PSPG", gamma=0.07)
    c.set_mesh(mesh)
    bad = elem.copy()
    bad[0] = (bad[0] + 3) % len(elem)  # the facet's nodes are not on that element
    with pytest.raises(pfem_amd.PfemError) as e:
        c.set_surface_facets(fac, bad)
    assert e.value.code == _lib.PFEM_E_INVALID
    c.close()
