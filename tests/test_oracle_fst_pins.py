"""Pins of the oracle's surface-tension term (getFST, MatricesBuilder.inl:389-402 with getP / getT
:167-214, and the facet loops of m_applyBCPSPG, MomContEquationPSPG.inl:155-186, and of the WC
m_applyBC, WCompNewton/MomEquation.inl:312-335) by what the mathematics fixes:

* -gamma int_Gamma P : grad N_i over a facet is minus gamma times the derivative of the facet's
  measure with respect to node i (P = I - n n, the tangential projector), so on a closed
  polygon / triangulated surface the nodal force is -gamma dA/dx_i: for a regular N-gon of
  radius R, 2 gamma sin(pi/N) towards the centre on every vertex, and on any polyhedron the
  closed form 1/2 sum (b - c) x n over the incident triangles; the element's node off the facet
  gets no force (grad N is normal to the facet there);
* the PSPG loop takes facets with any node on the free surface, the WC loop those with every
  node on it, and the WC step turns the force into a = FST / m_i with the lumped mass.
"""
import numpy as np
import pytest

from oracle import oracle as O
from pfem_amd import meshgen as mg

ONFS = 4


def polygon_drop(n=12, R=0.3, jitter=0.0, seed=0):
    rng = np.random.default_rng(seed)
    th = 2 * np.pi * np.arange(n) / n + jitter * rng.uniform(-0.2, 0.2, n) * 2 * np.pi / n
    r = R * (1 + jitter * rng.uniform(-0.2, 0.2, n))
    P = np.vstack([[0.0, 0.0], np.stack([r * np.cos(th), r * np.sin(th)], axis=1)])
    conn = np.array([[0, 1 + i, 1 + (i + 1) % n] for i in range(n)], dtype=np.int32)
    fac = np.array([[1 + i, 1 + (i + 1) % n] for i in range(n)], dtype=np.int32)
    flags = np.array([0] + [ONFS] * n, dtype=np.uint8)
    return mg.PfemMesh(2, P.T.copy(), conn, flags, R), fac, np.arange(n, dtype=np.int32)


def icosa_drop(R=0.3, jitter=0.0, seed=0):
    t = (1 + 5 ** 0.5) / 2
    V = np.array([[-1, t, 0], [1, t, 0], [-1, -t, 0], [1, -t, 0], [0, -1, t], [0, 1, t], [0, -1, -t], [0, 1, -t],
                  [t, 0, -1], [t, 0, 1], [-t, 0, -1], [-t, 0, 1]], dtype=float)
    V = R * V / np.linalg.norm(V, axis=1, keepdims=True)
    rng = np.random.default_rng(seed)
    V *= 1 + jitter * rng.uniform(-0.15, 0.15, (12, 1))
    F = np.array([[0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11], [1, 5, 9], [5, 11, 4], [11, 10, 2],
                  [10, 7, 6], [7, 1, 8], [3, 9, 4], [3, 4, 2], [3, 2, 6], [3, 6, 8], [3, 8, 9], [4, 9, 5],
                  [2, 4, 11], [6, 2, 10], [8, 6, 7], [9, 8, 1]])
    P = np.vstack([[0.0, 0.0, 0.0], V])
    conn = []
    for f in F:
        a, b, c = f + 1
        tet = [0, a, b, c]
        if np.linalg.det(np.stack([P[a], P[b], P[c]])) < 0:  # positive orientation with the centre first
            tet = [0, a, c, b]
        conn.append(tet)
    conn = np.array(conn, dtype=np.int32)
    fac = (F + 1).astype(np.int32)
    flags = np.array([0] + [ONFS] * 12, dtype=np.uint8)
    return mg.PfemMesh(3, P.T.copy(), conn, flags, R), fac, np.arange(len(F), dtype=np.int32)


def measure_gradient(mesh, fac):
    """d(sum of facet measures)/dx_i, closed form."""
    P = mesh.pos.T
    G = np.zeros_like(P)
    for f in fac:
        if mesh.dim == 2:
            a, b = P[f[0]], P[f[1]]
            t = (a - b) / np.linalg.norm(a - b)
            G[f[0]] += t
            G[f[1]] -= t
        else:
            a, b, c = P[f[0]], P[f[1]], P[f[2]]
            nrm = np.cross(b - a, c - a)
            nh = nrm / np.linalg.norm(nrm)
            G[f[0]] += 0.5 * np.cross(b - c, nh)
            G[f[1]] += 0.5 * np.cross(c - a, nh)
            G[f[2]] += 0.5 * np.cross(a - b, nh)
    return G.T


def pspg_fst(mesh, fac, elem, gamma):
    d, N = mesh.dim, mesh.n_nodes
    q = np.zeros((d + 1) * N)
    g = np.zeros(d)
    base = O.pspg_assemble(mesh, q, g=g)
    with O.surface_tension(fac, elem, gamma):
        st = O.pspg_assemble(mesh, q, g=g)
    assert np.array_equal(base.vals, st.vals)
    diff = st.b - base.b
    assert np.abs(diff[d * N:]).max() == 0.0  # continuity rows untouched
    return diff[: d * N].reshape(d, N)


def test_regular_polygon_vertex_force():
    n, R, gamma = 12, 0.3, 0.072
    mesh, fac, elem = polygon_drop(n, R)
    F = pspg_fst(mesh, fac, elem, gamma)
    P = mesh.pos
    want = -2 * gamma * np.sin(np.pi / n) * P[:, 1:] / R
    assert np.abs(F[:, 1:] - want).max() < 1e-14
    assert np.abs(F[:, 0]).max() < 1e-15


@pytest.mark.parametrize("dim", [2, 3])
def test_force_is_minus_gamma_measure_gradient(dim):
    gamma = 0.05
    mesh, fac, elem = polygon_drop(9, 0.2, jitter=1.0, seed=3) if dim == 2 else icosa_drop(0.2, jitter=1.0, seed=4)
    F = pspg_fst(mesh, fac, elem, gamma)
    want = -gamma * measure_gradient(mesh, fac)
    assert np.abs(F - want).max() < 1e-13 * gamma
    assert np.abs(F[:, 0]).max() < 1e-14 * gamma  # the interior node off every facet


def test_pspg_takes_facets_with_any_free_surface_node():
    gamma = 0.07
    mesh, fac, elem = polygon_drop(8, 0.25)
    mesh.flags[[1, 2]] = 0  # facet (1, 2) has no free-surface node; (8, 1) and (2, 3) have one
    F = pspg_fst(mesh, fac, elem, gamma)
    keep = [f for f in fac if (mesh.flags[f] & ONFS).any()]
    want = -gamma * measure_gradient(mesh, np.array(keep))
    assert np.abs(F - want).max() < 1e-14


@pytest.mark.parametrize("dim", [2, 3])
def test_wc_acceleration_is_fst_over_lumped_mass(dim):
    gamma = 0.06
    mesh, fac, elem = polygon_drop(10, 0.2, jitter=1.0, seed=5) if dim == 2 else icosa_drop(0.2, jitter=1.0, seed=6)
    if dim == 2:
        mesh.flags[3] = 0  # the two facets at node 3 lose their FST under the every-node rule
    N = mesh.n_nodes
    rho_star = 1000.0
    v = np.zeros(dim * N)
    a = np.zeros(dim * N)
    p = np.zeros(N)
    rho = np.full(N, rho_star)
    pos = mesh.pos.copy()
    with O.surface_tension(fac, elem, gamma):
        O.wc_step(mesh, pos, v, p, rho, a, 1e-9, mu=0.0, g=[0.0] * dim, rho_star=rho_star)
    keep = [f for f in fac if (mesh.flags[f] & ONFS).all()]
    Fst = -gamma * measure_gradient(mesh, np.array(keep))
    # lumped mass rho sum_e |e| / (d+1)
    P = mesh.pos.T
    m = np.zeros(N)
    for e in mesh.conn:
        X = P[e]
        vol = abs(np.linalg.det((X[1:] - X[0]).T)) / (2 if dim == 2 else 6)
        m[e] += rho_star * vol / (dim + 1)
    want = Fst / m
    assert np.abs(a.reshape(dim, N) - want).max() < 1e-10 * np.abs(want).max()
