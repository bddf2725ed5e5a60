"""Pins of the oracle's heat equation (HeatEqIncompNewton without phase change,
srcs/simulation/physics/IncompNewton/HeatEquation.inl) by what the mathematics fixes:

* energy balance: with no Dirichlet node, the rows of (M + dt L) T = M T0 - dt (facet terms)
  sum to  sum_i m_i (T_i - T0) = -dt [h (T0 - Tinf) + sigma eps (T0^4 - Tinf^4)] |Gamma|  for a
  uniform T0 (L's rows and columns sum to zero, the lumped masses m_i and the boundary measure
  computed here from the mesh);
* facet terms on hand values: a constant flux gives (q . n_out) |f| / d per node with the
  outward normal, a linear temperature along a segment gives h (2 T_a + T_b) |f| / 6, radiation
  sigma eps (T^4 - Tinf^4) |f| / d for a uniform T;
* a linear field held on the boundary stays the solution (the discrete Laplacian reproduces
  linear fields and M (T - Tprev) vanishes);
* the step equals the reduced solve of the system assembled here from the P1 closed forms
  M_ij = rho cv |e| (1 + d_ij) / ((d+1)(d+2)), L_ij = k |e| gradNi . gradNj with the Dirichlet
  nodes eliminated;
* a node in no element keeps its temperature, a Dirichlet node takes its value;
* the oracle's CG solution satisfies its own system to the requested tolerance and equals a
  direct solve.
"""
import itertools

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spl

import pfem_amd
from oracle import oracle as O
from pfem_amd import meshgen as mg

SIGMA = 5.670374419e-8


def block_mesh(dim, n=6, h=0.1, seed=0, walls=False):
    """A jittered block of nodes triangulated by the mesher (facets = its boundary)."""
    rng = np.random.default_rng(seed)
    P = np.array(list(itertools.product(*[range(n)] * dim)), dtype=float) * h
    inner = ((P > 0) & (P < (n - 1) * h - 1e-12)).all(axis=1)
    P[inner] += rng.uniform(-0.15 * h, 0.15 * h, size=(inner.sum(), dim))
    flags = np.zeros(len(P), np.uint8)
    m = pfem_amd.Mesher(dim, hchar=h, alpha=1.3)
    m.set_nodes(P.T.copy(), flags)
    m.triangulate()
    pos, _, fl, _ = m.nodes()
    mesh = mg.PfemMesh(dim, pos, m.elements(), fl, h)
    fac, out, _ = m.facets()
    return mesh, fac, out


def element_geo(mesh):
    P = mesh.pos.T
    E = mesh.conn
    d = mesh.dim
    X = P[E]                       # E x (d+1) x d
    J = (X[:, 1:] - X[:, :1]).transpose(0, 2, 1)  # columns = edges
    det = np.linalg.det(J)
    vol = det / (2 if d == 2 else 6)
    invJ = np.linalg.inv(J)
    G = np.zeros((len(E), d, d + 1))
    G[:, :, 1:] = invJ.transpose(0, 2, 1)
    G[:, :, 0] = -G[:, :, 1:].sum(axis=2)
    return vol, G


def facet_measure(mesh, fac):
    P = mesh.pos.T
    if mesh.dim == 2:
        return np.linalg.norm(P[fac[:, 1]] - P[fac[:, 0]], axis=1)
    return 0.5 * np.linalg.norm(np.cross(P[fac[:, 1]] - P[fac[:, 0]], P[fac[:, 2]] - P[fac[:, 0]]), axis=1)


@pytest.mark.parametrize("dim", [2, 3])
def test_heat_energy_balance(dim):
    mesh, fac, out = block_mesh(dim, n=6 if dim == 2 else 4)
    hc = O.HeatCase(rho=1000.0, cv=4000.0, k=0.6, h=25.0, Tinf=280.0, epsRad=0.8, facets=fac, facet_out=out,
                    facet_kind=np.full(len(fac), O.HEAT_QH | O.HEAT_QR, np.uint8))
    T0 = 350.0
    T = np.full(mesh.n_nodes, T0)
    dt = 5.0
    st, it, rr = O.heat_step(mesh, T, hc=hc, dt=dt, rtol=1e-15)
    assert st in (0, 1) and rr < 1e-13
    vol, _ = element_geo(mesh)
    m = np.zeros(mesh.n_nodes)
    np.add.at(m, mesh.conn.reshape(-1), np.repeat(vol / (dim + 1), dim + 1))
    m *= hc.rho * hc.cv
    gamma = facet_measure(mesh, fac).sum()
    expect = -dt * (hc.h * (T0 - hc.Tinf) + SIGMA * hc.epsRad * (T0 ** 4 - hc.Tinf ** 4)) * gamma
    got = float((m * (T - T0)).sum())
    assert got == pytest.approx(expect, rel=1e-9)
    # the boundary cools (the consistent mass matrix lets single interior nodes overshoot T0)
    assert T[np.unique(fac)].mean() < T0


def test_heat_facet_terms_on_hand_values():
    # 2D: one triangle (0,0) (1,0) (0,1); facet (0,0)-(1,0) with out node (0,1): outward normal -y
    mesh = mg.PfemMesh(2, np.array([[0.0, 2.0, 0.0], [0.0, 0.0, 1.0]]), np.array([[0, 1, 2]], np.int32),
                       np.zeros(3, np.uint8), 1.0)
    q = np.array([0.3, -1.5])
    hc = O.HeatCase(h=7.0, Tinf=10.0, epsRad=0.5, facets=np.array([[0, 1]]), facet_out=np.array([2]),
                    facet_kind=np.array([7], np.uint8), qgp=np.tile(q, (1, 3, 1)))
    T = np.array([20.0, 50.0, 0.0])
    qn, sgh, sgr = O.heat_facet_terms(mesh, T, hc)
    L = 2.0
    np.testing.assert_allclose(qn[0], [(q @ [0, -1]) * L / 2] * 2, rtol=1e-14)
    # linear T along the segment: int N_0 h (T - Tinf) = h L (2 Ta + Tb) / 6 - h Tinf L / 2
    Ta, Tb = 20.0, 50.0
    np.testing.assert_allclose(sgh[0], [7.0 * L * (2 * Ta + Tb) / 6 - 7.0 * 10.0 * L / 2,
                                         7.0 * L * (Ta + 2 * Tb) / 6 - 7.0 * 10.0 * L / 2], rtol=1e-14)
    # radiation with a uniform facet temperature
    Tu = np.array([40.0, 40.0, 0.0])
    _, _, sgr = O.heat_facet_terms(mesh, Tu, hc)
    np.testing.assert_allclose(sgr[0], [SIGMA * 0.5 * (40.0 ** 4 - 10.0 ** 4) * L / 2] * 2, rtol=1e-13)
    # 3D: the facet z = 0 of the unit tet, out node (0,0,1): outward normal -z, area 1/2
    mesh3 = mg.PfemMesh(3, np.array([[0.0, 1, 0, 0], [0, 0, 1, 0], [0, 0, 0, 1]]), np.array([[0, 1, 2, 3]], np.int32),
                        np.zeros(4, np.uint8), 1.0)
    q3 = np.array([0.2, 0.1, 2.0])
    hc3 = O.HeatCase(h=3.0, Tinf=0.0, facets=np.array([[0, 2, 1]]), facet_out=np.array([3]),
                     facet_kind=np.array([3], np.uint8), qgp=np.tile(q3, (1, 3, 1)))
    qn3, sgh3, _ = O.heat_facet_terms(mesh3, np.full(4, 5.0), hc3)
    np.testing.assert_allclose(qn3[0], [(q3 @ [0, 0, -1]) * 0.5 / 3] * 3, rtol=1e-14)
    np.testing.assert_allclose(sgh3[0], [3.0 * 5.0 * 0.5 / 3] * 3, rtol=1e-14)


@pytest.mark.parametrize("dim", [2, 3])
def test_heat_linear_field_held_on_the_boundary_is_steady(dim):
    mesh, fac, out = block_mesh(dim, n=6 if dim == 2 else 4, seed=3)
    a = np.array([3.0, -2.0, 1.5][:dim])
    lin = a @ mesh.pos + 300.0
    fl = mesh.flags.copy()
    fl[np.unique(fac)] |= mg.DIRICHLET
    m2 = mg.PfemMesh(dim, mesh.pos, mesh.conn, fl, mesh.h)
    T = lin.copy()
    st, _, _ = O.heat_step(m2, T, Tbc=lin, hc=O.HeatCase(), dt=10.0, rtol=1e-15)
    assert st in (0, 1)
    np.testing.assert_allclose(T, lin, rtol=0, atol=1e-10 * np.abs(lin).max())


@pytest.mark.parametrize("dim", [2, 3])
def test_heat_step_equals_reduced_closed_form_solve(dim):
    mesh, fac, out = block_mesh(dim, n=6 if dim == 2 else 4, seed=5)
    N = mesh.n_nodes
    rng = np.random.default_rng(1)
    T0 = 300.0 + rng.normal(scale=5.0, size=N)
    fl = mesh.flags.copy()
    bnd = np.unique(fac)
    dir_nodes = bnd[mesh.pos[0, bnd] < 1e-12]          # the x = 0 side held
    fl[dir_nodes] |= mg.DIRICHLET
    Tbc = np.zeros(N)
    Tbc[dir_nodes] = 400.0 + 10 * mesh.pos[1, dir_nodes]
    hc = O.HeatCase(rho=900.0, cv=2000.0, k=3.0)
    dt = 2.0
    m2 = mg.PfemMesh(dim, mesh.pos, mesh.conn, fl, mesh.h)
    T = T0.copy()
    st, _, _ = O.heat_step(m2, T, Tbc=Tbc, hc=hc, dt=dt, rtol=1e-15)
    assert st in (0, 1)
    vol, G = element_geo(mesh)
    ne = dim + 1
    Mloc = (np.ones((ne, ne)) + np.eye(ne)) / ((dim + 1) * (dim + 2))
    rows, cols, vals = [], [], []
    for e, nodes in enumerate(mesh.conn):
        Me = hc.rho * hc.cv * vol[e] * Mloc
        Le = hc.k * vol[e] * G[e].T @ G[e]
        for i in range(ne):
            for j in range(ne):
                rows.append(nodes[i])
                cols.append(nodes[j])
                vals.append(Me[i, j] + dt * Le[i, j])
    A = sp.csr_matrix((vals, (rows, cols)), shape=(N, N))
    Mm = sp.csr_matrix(([hc.rho * hc.cv * vol[e] * Mloc[i, j] for e in range(len(vol)) for i in range(ne)
                         for j in range(ne)], ([n for nodes in mesh.conn for n in np.repeat(nodes, ne)],
                                               [n for nodes in mesh.conn for _ in range(ne) for n in nodes])),
                       shape=(N, N))
    b = Mm @ T0
    free = np.setdiff1d(np.arange(N), dir_nodes)
    x = np.zeros(N)
    x[dir_nodes] = Tbc[dir_nodes]
    rhs = b[free] - A[free][:, dir_nodes] @ x[dir_nodes]
    x[free] = spl.spsolve(A[free][:, free].tocsc(), rhs)
    np.testing.assert_allclose(T, x, rtol=1e-11, atol=0)


def test_heat_free_node_keeps_its_temperature_and_cg_meets_its_tolerance():
    mesh, fac, out = block_mesh(2, n=5, seed=7)
    pos = np.hstack([mesh.pos, [[5.0], [5.0]]])          # a node in no element
    fl = np.concatenate([mesh.flags, [0]]).astype(np.uint8)
    m2 = mg.PfemMesh(2, pos, mesh.conn, fl, mesh.h)
    rng = np.random.default_rng(2)
    T = 300 + rng.normal(size=m2.n_nodes)
    T_free = T[-1]
    hc = O.HeatCase(h=10.0, Tinf=250.0, facets=fac, facet_out=out, facet_kind=np.full(len(fac), O.HEAT_QH, np.uint8))
    S = O.heat_assemble(m2, T, hc=hc, dt=1.0)
    A = sp.csr_matrix((S.vals, S.col_idx, S.row_ptr), shape=(S.n, S.n))
    # the free node's row is the unit row with its own temperature as rhs
    assert A[S.n - 1].nnz == 1 and A[S.n - 1, S.n - 1] == 1.0 and S.b[-1] == T_free
    assert (A - A.T).count_nonzero() == 0 or abs(A - A.T).max() <= 1e-12 * abs(A).max()
    x = T.copy()
    st, it, rr = O.heat_step(m2, x, hc=hc, dt=1.0, rtol=1e-12)
    assert st == 0 and rr <= 1e-12 and it > 0
    assert np.linalg.norm(A @ x - S.b) <= 1e-12 * np.linalg.norm(S.b)
    np.testing.assert_allclose(x, spl.spsolve(A.tocsc(), S.b), rtol=1e-10)
    assert x[-1] == T_free
