"""Pins of the oracle's Boussinesq factors (MomContEquation.inl:164-197: F factor
rho (1 - alpha (T - Tr)), PSPG H factor 1 - alpha (T - Tr), T = N T_e at each Gauss point) by what
the mathematics fixes:

* a uniform temperature Tr + dT is a body force scaled by 1 - alpha dT: the Boussinesq step equals
  the Newtonian step with g (1 - alpha dT);
* for a temperature field, the right-hand side moves by the P1 integrals of the linear factor,
  element by element: momentum rows by -rho alpha g_s int N_a (T - Tr) = -rho alpha g_s |e|
  ((T_a + sum_c T_c) / ((d+1)(d+2)) - Tr / (d+1)), continuity rows by
  -alpha tau_e |e| (gradN_a . g) (mean_c T_c - Tr), with tau_e the element's PSPG parameter;
* T = Tr leaves the system of the Newtonian fluid unchanged (bit for bit in 2D).
"""
import numpy as np
import pytest

from oracle import oracle as O
from test_oracle_heat_pins import block_mesh, element_geo


def gvec(dim):
    return np.array([0.0, -9.81] if dim == 2 else [0.0, 0.0, -9.81])


def state(mesh, seed):
    rng = np.random.default_rng(seed)
    d, N = mesh.dim, mesh.n_nodes
    v = 0.05 * rng.standard_normal((d, N))
    p = 1000.0 * (mesh.pos.T @ gvec(d)) + rng.standard_normal(N)
    return np.concatenate([v.ravel(), p])


@pytest.mark.parametrize("dim", [2, 3])
def test_uniform_temperature_scales_gravity(dim):
    mesh, _, _ = block_mesh(dim, n=7 if dim == 2 else 5, seed=4)
    N = mesh.n_nodes
    P = mesh.pos.T
    hi = P.max(axis=0)
    wall = ((P < 1e-12) | (P > hi - 1e-12)).any(axis=1)
    mesh.flags = np.where(wall, 1 | 8 | 16, 0).astype(np.uint8)
    alpha, Tr, dT = 2e-3, 300.0, 40.0
    q0 = state(mesh, 1)
    g = gvec(dim)
    qa, pa = q0.copy(), mesh.pos.copy()
    with O.material("Boussinesq", alpha=alpha, Tr=Tr, T=np.full(N, Tr + dT)):
        sa, _ = O.pspg_step(mesh, pa, qa, g=g, dt=2e-3, residual="U", min_res=1e-9, rtol=1e-13)
    qb, pb = q0.copy(), mesh.pos.copy()
    sb, _ = O.pspg_step(mesh, pb, qb, g=g * (1 - alpha * dT), dt=2e-3, residual="U", min_res=1e-9, rtol=1e-13)
    assert sa == 0 and sb == 0
    assert np.abs(qa - qb).max() < 1e-9 * np.abs(qb).max()
    assert np.abs(pa - pb).max() < 1e-13


@pytest.mark.parametrize("dim", [2, 3])
def test_temperature_field_moves_the_rhs_by_p1_integrals(dim):
    mesh, _, _ = block_mesh(dim, n=6 if dim == 2 else 4, seed=5)
    mesh.flags = np.zeros(mesh.n_nodes, np.uint8)
    d, N = dim, mesh.n_nodes
    rng = np.random.default_rng(9)
    T = 300.0 + 50.0 * rng.random(N)
    alpha, Tr, rho, mu, dt = 3e-3, 320.0, 1000.0, 1e-3, 1e-3
    g = gvec(dim)
    q = np.zeros((d + 1) * N)
    base = O.pspg_assemble(mesh, q, rho=rho, mu=mu, g=g, dt=dt, apply_bc=False)
    with O.material("Boussinesq", alpha=alpha, Tr=Tr, T=T):
        bq = O.pspg_assemble(mesh, q, rho=rho, mu=mu, g=g, dt=dt, apply_bc=False)
    assert np.array_equal(base.vals, bq.vals)
    vol, G = element_geo(mesh)
    ne = d + 1
    want = np.zeros((d + 1) * N)
    for e, nodes in enumerate(mesh.conn):
        Te = T[nodes]
        pos_e = mesh.pos[:, nodes]
        _, _, tau = O.pspg_element(d, pos_e, np.zeros((d + 1) * ne), np.zeros((d + 1) * ne), rho=rho, mu=mu, g=g,
                                   dt=dt)
        for a, n in enumerate(nodes):
            intT = vol[e] * ((Te[a] + Te.sum()) / (ne * (ne + 1)) - Tr / ne)
            for s in range(d):
                want[n + s * N] += -rho * alpha * g[s] * intT
            want[n + d * N] += -alpha * tau * vol[e] * (G[e, :, a] @ g) * (Te.mean() - Tr)
    diff = bq.b - base.b
    assert np.abs(diff - want).max() < 1e-10 * np.abs(want).max()


@pytest.mark.parametrize("dim", [2, 3])
def test_reference_temperature_is_newtonian(dim):
    mesh, _, _ = block_mesh(dim, n=6 if dim == 2 else 4, seed=6)
    N = mesh.n_nodes
    q = state(mesh, 2)
    base = O.pspg_assemble(mesh, q, g=gvec(dim))
    with O.material("Boussinesq", alpha=5e-3, Tr=300.0, T=np.full(N, 300.0)):
        bq = O.pspg_assemble(mesh, q, g=gvec(dim))
    assert np.array_equal(base.vals, bq.vals)
    if dim == 2:
        assert np.array_equal(base.b, bq.b)
    else:  # the 3D Gauss points are given to 15 digits: sum_a N_a = 1 to rounding only
        assert np.abs(base.b - bq.b).max() < 1e-14 * np.abs(base.b).max()
