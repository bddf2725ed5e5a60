"""Pins of the CPU oracle (oracle/pfem_oracle.cpp) against what the mathematics fixes.

The reference ships no tests, golden vectors or fixtures for this path, so the oracle is
pinned by independent closed forms of the P1 element integrals, invariants, brute-force
pattern/value checks on 1-2 element meshes, a SciPy SuperLU golden solution
(tests/golden/), conservation properties of the explicit step and the CFL closed form.
Each pin is chosen so that a plausible restatement mistake (a dropped Gauss weight, a
transposed gradient, a wrong sign or index) fails it.
"""
import itertools
import math
import os

import numpy as np
import pytest

from oracle import oracle as O
from pfem_amd import meshgen as mg

RNG = np.random.default_rng(1234)


def rand_simplex(dim, scale=0.1):
    """A positively oriented, well-shaped random simplex (dim x (dim+1), field-blocked)."""
    while True:
        P = RNG.uniform(-1, 1, size=(dim, dim + 1)) * scale
        J = P[:, 1:] - P[:, :1]
        dj = np.linalg.det(J)
        if dj > 0.05 * scale ** dim:
            return P


def exact_gradN(P):
    """Gradients of the P1 barycentric functions, from the linear system [1 x; ...] c = e_i."""
    d = P.shape[0]
    A = np.vstack([np.ones(d + 1), P]).T  # (d+1) x (d+1): rows [1, x_i]
    C = np.linalg.inv(A)  # column i: coefficients of N_i = c0 + c.x
    return C[1:, :]  # d x (d+1)


def volume(P):
    d = P.shape[0]
    return abs(np.linalg.det(P[:, 1:] - P[:, :1])) / math.factorial(d)


@pytest.mark.parametrize("dim", [2, 3])
def test_element_matrices_closed_forms(dim):
    for _ in range(5):
        P = rand_simplex(dim)
        rho, mu = 997.0, 1.7e-3
        m = O.element_matrices(dim, P, rho=rho, mu=mu)
        G = exact_gradN(P)
        vol = volume(P)
        ne = dim + 1
        # detJ = d! |e|
        assert m["detJ"] == pytest.approx(math.factorial(dim) * vol, rel=1e-13)
        # gradN (MatricesBuilder.inl:93-128) vs the barycentric gradients
        np.testing.assert_allclose(m["gradN"], G, rtol=0, atol=1e-12 * np.abs(G).max())
        # M = rho |e| (1 + delta_ij) / ((d+1)(d+2)): exact P1 mass matrix
        Mex = rho * vol * (np.ones((ne, ne)) + np.eye(ne)) / ((dim + 1) * (dim + 2))
        np.testing.assert_allclose(m["M"], Mex, rtol=1e-13)
        # K_(is)(jt) = mu |e| (delta_st gradNi.gradNj + d_t N_i d_s N_j)  (tensor form of 2 mu sym(grad))
        Kex = np.zeros((dim * ne, dim * ne))
        for s, t in itertools.product(range(dim), repeat=2):
            for i, j in itertools.product(range(ne), repeat=2):
                Kex[i + s * ne, j + t * ne] = mu * vol * ((s == t) * G[:, i] @ G[:, j] + G[t, i] * G[s, j])
        np.testing.assert_allclose(m["K"], Kex, rtol=0, atol=1e-12 * np.abs(Kex).max())
        # D_(i)(j,k) = |e|/(d+1) d_k N_j ; C_(i)(j,k) = |e|/(d+1) d_k N_i ; L = |e| gradN^T gradN
        Dex = np.zeros((ne, dim * ne))
        Cex = np.zeros((ne, dim * ne))
        for i, j, k in itertools.product(range(ne), range(ne), range(dim)):
            Dex[i, j + k * ne] = vol / ne * G[k, j]
            Cex[i, j + k * ne] = vol / ne * G[k, i]
        np.testing.assert_allclose(m["D"], Dex, rtol=0, atol=1e-12 * np.abs(Dex).max())
        np.testing.assert_allclose(m["C"], Cex, rtol=0, atol=1e-12 * np.abs(Cex).max())
        np.testing.assert_allclose(m["L"], vol * G.T @ G, rtol=0, atol=1e-12 * np.abs(G.T @ G).max() * vol)


@pytest.mark.parametrize("dim", [2, 3])
def test_element_invariants(dim):
    P = rand_simplex(dim)
    m = O.element_matrices(dim, P, rho=1.0, mu=1.0)
    ne = dim + 1
    G = m["gradN"]
    # sum_i gradN_i = 0 and sum_i x_i (x) gradN_i = I
    np.testing.assert_allclose(G.sum(axis=1), 0, atol=1e-10 * np.abs(G).max())
    np.testing.assert_allclose(P @ G.T, np.eye(dim), atol=1e-12)
    K = m["K"]
    np.testing.assert_allclose(K, K.T, atol=1e-12 * np.abs(K).max())
    assert np.linalg.eigvalsh(K).min() > -1e-10 * np.abs(K).max()
    # rigid translations and rotations have zero strain rate
    for s in range(dim):
        v = np.zeros(dim * ne)
        v[s * ne:(s + 1) * ne] = 1.0
        np.testing.assert_allclose(K @ v, 0, atol=1e-10 * np.abs(K).max())
        np.testing.assert_allclose(m["D"] @ v, 0, atol=1e-10 * np.abs(m["D"]).max())
    for a, b in itertools.combinations(range(dim), 2):
        v = np.zeros(dim * ne)
        v[a * ne:(a + 1) * ne] = -P[b]
        v[b * ne:(b + 1) * ne] = P[a]
        np.testing.assert_allclose(K @ v, 0, atol=1e-9 * np.abs(K).max())
    np.testing.assert_allclose(m["L"] @ np.ones(ne), 0, atol=1e-10 * np.abs(m["L"]).max())


def test_inradius_closed_forms():
    # right triangle with legs 1: r = (a + b - c)/2
    P2 = np.array([[0.0, 1.0, 0.0], [0.0, 0.0, 1.0]])
    assert O.element_matrices(2, P2)["rin"] == pytest.approx((2 - math.sqrt(2)) / 2, rel=1e-14)
    # trirectangular tet: r = 3V / A_total
    P3 = np.array([[0.0, 1.0, 0.0, 0.0], [0.0, 0.0, 1.0, 0.0], [0.0, 0.0, 0.0, 1.0]])
    A = 3 * 0.5 + math.sqrt(3) / 2
    assert O.element_matrices(3, P3)["rin"] == pytest.approx(3 * (1 / 6) / A, rel=1e-14)


@pytest.mark.parametrize("dim", [2, 3])
def test_pspg_element_closed_form(dim):
    """Ae, be of MomContEquationPSPG.inl:42,53 from the exact P1 integrals."""
    ne = dim + 1
    rho, mu, dt = 1000.0, 1e-3, 1e-3
    g = np.array([0.0, -9.81] if dim == 2 else [0.0, 0.0, -9.81])
    P = rand_simplex(dim, scale=0.05)
    qprev = RNG.normal(size=(dim + 1) * ne)
    qiter = RNG.normal(size=(dim + 1) * ne)
    Ae, be, tau = O.pspg_element(dim, P, qprev, qiter, rho=rho, mu=mu, g=list(g), dt=dt)
    vol = volume(P)
    G = exact_gradN(P)
    # tau: MomContEquationPSPG.inl:238-259
    h = math.sqrt((0.5 if dim == 2 else 1 / 6) * math.factorial(dim) * vol / math.pi)
    U = np.mean([np.linalg.norm(qiter[[i + s * ne for s in range(dim)]]) for i in range(ne)])
    tau_ex = 1 / math.sqrt((2 / dt) ** 2 + (2 * U / h) ** 2 + 9 * (4 * mu / (h * h * rho)) ** 2)
    assert tau == pytest.approx(tau_ex, rel=1e-13)
    M = rho * vol * (np.ones((ne, ne)) + np.eye(ne)) / ((dim + 1) * (dim + 2))
    na = (dim + 1) * ne
    A = np.zeros((na, na))
    for s, t in itertools.product(range(dim), repeat=2):
        for i, j in itertools.product(range(ne), repeat=2):
            A[i + s * ne, j + t * ne] = (s == t) * M[i, j] / dt + mu * vol * ((s == t) * G[:, i] @ G[:, j] + G[t, i] * G[s, j])
    for s in range(dim):
        for i, j in itertools.product(range(ne), repeat=2):
            A[i + s * ne, j + dim * ne] = -vol / ne * G[s, i]
            A[i + dim * ne, j + s * ne] = vol / ne * (tau / dt * G[s, i] + G[s, j])
    for i, j in itertools.product(range(ne), repeat=2):
        A[i + dim * ne, j + dim * ne] = tau / rho * vol * G[:, i] @ G[:, j]
    np.testing.assert_allclose(Ae, A, rtol=0, atol=1e-12 * np.abs(A).max())
    vp = qprev[:dim * ne]
    b = np.zeros(na)
    for s in range(dim):
        for i in range(ne):
            b[i + s * ne] = rho * vol * g[s] / ne + M[i] @ vp[s * ne:(s + 1) * ne] / dt
    for i in range(ne):
        b[i + dim * ne] = tau * vol * G[:, i] @ g + tau / dt * vol / ne * sum(
            G[s, i] * vp[s * ne:(s + 1) * ne].sum() for s in range(dim))
    np.testing.assert_allclose(be, b, rtol=0, atol=1e-12 * np.abs(b).max())


def two_element_mesh(dim, bound_nodes=(), free_extra=True, dirichlet=()):
    """Two simplices sharing a facet (+ one free node that belongs to no element)."""
    if dim == 2:
        pos = np.array([[0.0, 1.0, 0.0, 1.1, 3.0], [0.0, 0.0, 1.0, 1.05, 3.0]])
        conn = np.array([[0, 1, 2], [1, 3, 2]], np.int32)
    else:
        pos = np.array([[0.0, 1.0, 0.0, 0.0, 0.9, 3.0], [0.0, 0.0, 1.0, 0.0, 0.8, 3.0],
                        [0.0, 0.0, 0.0, 1.0, 0.7, 3.0]])
        conn = np.array([[0, 1, 2, 3], [1, 4, 2, 3]], np.int32)
    if not free_extra:
        pos = pos[:, :-1]
    N = pos.shape[1]
    flags = np.zeros(N, np.uint8)
    for b in bound_nodes:
        flags[b] |= mg.BOUND | mg.FIXED
    for b in dirichlet:
        flags[b] |= mg.BOUND | mg.DIRICHLET | mg.FIXED
    m = mg.PfemMesh(dim, pos, conn, flags, 1.0)
    m.surface = np.full(N, 2.0)
    return m


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("bound", [(), (0,), (1, 2)])
def test_pattern_and_values_brute_force(dim, bound):
    """setFromTriplets over the reference's triplet rules (MomContEquationPSPG.inl:58-128) == dense scatter."""
    m = two_element_mesh(dim, bound_nodes=bound)
    N = m.n_nodes
    ne = dim + 1
    n = (dim + 1) * N
    free = m.free_mask()
    qprev = RNG.normal(size=n)
    qiter = RNG.normal(size=n)
    S = O.pspg_assemble(m, qprev, qiter, dt=1e-3, apply_bc=False)
    A = S.to_scipy().toarray()
    exp = np.zeros((n, n))
    pattern = set()
    bexp = np.zeros(n)
    for e in range(m.n_elems):
        nodes = m.conn[e]
        Ae, be, _ = O.pspg_element(dim, m.pos[:, nodes], qprev.reshape(dim + 1, N)[:, nodes].reshape(-1),
                                   qiter.reshape(dim + 1, N)[:, nodes].reshape(-1), dt=1e-3)
        for a, b in itertools.product(range(ne), repeat=2):
            i, j = nodes[a], nodes[b]
            for s, t in itertools.product(range(dim + 1), repeat=2):
                row_emitted = (s == dim and not free[i]) or (s < dim and i not in bound and not free[i])
                if row_emitted:
                    exp[i + s * N, j + t * N] += Ae[a + s * ne, b + t * ne]
                    pattern.add((i + s * N, j + t * N))
        for a in range(ne):
            for s in range(dim + 1):
                bexp[nodes[a] + s * N] += be[a + s * ne]
    for i in range(N):
        if free[i] or i in bound:
            for s in range(dim):
                exp[i + s * N, i + s * N] = 1.0
                pattern.add((i + s * N, i + s * N))
        if free[i]:
            exp[i + dim * N, i + dim * N] = 1.0
            pattern.add((i + dim * N, i + dim * N))
    pattern.add((0, 0))  # skipped triplets default to (0, 0, 0)
    got = {(r, c) for r in range(n) for c in S.col_idx[S.row_ptr[r]:S.row_ptr[r + 1]]}
    assert got == pattern
    np.testing.assert_allclose(A, exp, rtol=0, atol=1e-12 * np.abs(exp).max())
    np.testing.assert_allclose(S.b, bexp, rtol=0, atol=1e-12 * np.abs(bexp).max())
    # CSR rows sorted
    for r in range(n):
        cols = S.col_idx[S.row_ptr[r]:S.row_ptr[r + 1]]
        assert np.all(np.diff(cols) > 0)


@pytest.mark.parametrize("dim", [2, 3])
def test_dirichlet_elimination_equivalence(dim):
    """Column elimination (MomContEquationPSPG.inl:216-228) keeps the solution: eliminated and
    un-eliminated systems (Dirichlet rows are identities with b = v_bc) solve to the same q."""
    m = two_element_mesh(dim, dirichlet=(0, 2))
    N = m.n_nodes
    n = (dim + 1) * N
    qprev = RNG.normal(size=n)
    dv = RNG.normal(size=dim * N)
    S0 = O.pspg_assemble(m, qprev, qprev, dir_vals=dv, dt=1e-3, apply_bc=False)
    S1 = O.pspg_assemble(m, qprev, qprev, dir_vals=dv, dt=1e-3, apply_bc=True)
    # pattern is unchanged (explicit zeros are kept)
    np.testing.assert_array_equal(S0.row_ptr, S1.row_ptr)
    np.testing.assert_array_equal(S0.col_idx, S1.col_idx)
    # un-eliminated system with the Dirichlet/free right-hand side set by hand
    b0 = S0.b.copy()
    free = m.free_mask()
    for i in range(N):
        if free[i]:
            b0[i + dim * N] = 0
            if not m.flags[i] & mg.BOUND:
                for s in range(dim):
                    b0[i + s * N] = qprev[i + s * N] + 1e-3 * ([0, -9.81] if dim == 2 else [0, 0, -9.81])[s]
        if m.flags[i] & mg.DIRICHLET:
            for s in range(dim):
                b0[i + s * N] = dv[i + s * N]
    x0 = np.linalg.solve(S0.to_scipy().toarray(), b0)
    x1 = np.linalg.solve(S1.to_scipy().toarray(), S1.b)
    np.testing.assert_allclose(x1, x0, rtol=0, atol=1e-9 * np.abs(x0).max())
    # eliminated columns are zero off the diagonal
    A1 = S1.to_scipy().toarray()
    for i in (0, 2):
        for s in range(dim):
            col = A1[:, i + s * N].copy()
            col[i + s * N] = 0
            assert np.all(col == 0)


def test_superlu_golden_c1():
    """The oracle's C1 system solved by the Krylov stand-in matches the committed SuperLU solution."""
    path = os.path.join(os.path.dirname(__file__), "golden", "c1_superlu.npz")
    gold = np.load(path)
    m = mg.dambreak_2d()
    q = mg.pspg_hydrostatic(m)
    S = O.pspg_assemble(m, q, dt=1e-3)
    assert S.n == int(gold["n"]) and len(S.vals) == int(gold["nnz"])
    np.testing.assert_array_equal(S.row_ptr, gold["row_ptr"])
    assert float(np.linalg.norm(S.b)) == pytest.approx(float(gold["b_norm"]), rel=1e-13)
    st, x, it, rr = O.solve(S, kind=2, rtol=1e-12)
    assert st == 0 and rr <= 1e-12
    xg = gold["x"]
    assert np.linalg.norm(x - xg) / np.linalg.norm(xg) < 1e-9


def test_wc_rigid_translation_fixed_point():
    """No gravity, uniform pressure and a rigid translation: a = 0, p and rho unchanged,
    positions shift by dt*v exactly (div v = 0, K v = 0, D^T P = 0 for uniform P on interior)."""
    m = mg.tank_2d(h=0.1, size=(0.6, 0.4), fill=1.0)
    m.flags[:] = 0  # no walls: every node moves
    N = m.n_nodes
    pos = m.pos.reshape(-1).copy()
    v = np.zeros(2 * N)
    v[:N] = 0.3
    v[N:] = -0.2
    p = np.full(N, 0.0)
    rho = np.full(N, 1000.0)
    a = np.zeros(2 * N)
    pos0 = pos.copy()
    dt = 1e-4
    O.wc_step(m, pos, v, p, rho, a, dt, g=[0.0, 0.0], K0=2.2e5, K0p=7.6, rho_star=1000.0)
    np.testing.assert_allclose(a, 0, atol=1e-9)
    np.testing.assert_allclose(p, 0, atol=1e-6)
    np.testing.assert_allclose(rho, 1000.0, rtol=1e-14)
    np.testing.assert_allclose(v[:N], 0.3, rtol=1e-12)
    np.testing.assert_allclose(pos[:N], pos0[:N] + 0.3 * dt, rtol=0, atol=1e-15)


def test_wc_tait_roundtrip_and_cfl_closed_form():
    """EOS p -> rho of ContEquation.inl:319-332 and the CFL dt of WCompNewton/Solver.cpp:192-234
    on an un-jittered lattice of right triangles at rest: dt = sc * he / c with he = 2 r_in."""
    h = 0.1
    m = mg.tank_2d(h=h, size=(0.4, 0.2), fill=1.0)
    m.pos = np.round(m.pos / h) * h  # remove jitter
    N = m.n_nodes
    K0, K0p, rs = 2.2e5, 7.6, 1000.0
    p = np.zeros(N)
    rho = np.full(N, rs)
    v = np.zeros(2 * N)
    dt = O.wc_next_dt(m, m.pos.reshape(-1).copy(), v, p, rho, mu=0.0, K0=K0, K0p=K0p, security_coeff=0.5,
                      max_dt=1.0)
    rin = h * (2 - math.sqrt(2)) / 2
    c = math.sqrt(K0 / rs)
    assert dt == pytest.approx(0.5 * 2 * rin / c, rel=1e-12)
    # Tait round trip: rho(p) then p(rho) (the reference's two EOS directions)
    pp = np.array([0.0, 1e3, 5e4])
    r = rs * ((K0p / K0) * pp + 1) ** (1 / K0p)
    np.testing.assert_allclose(K0 / K0p * ((r / rs) ** K0p - 1), pp, rtol=1e-10, atol=1e-9)


def test_picard_step_c1_converges():
    m = mg.dambreak_2d()
    q = mg.pspg_hydrostatic(m)
    pos = m.pos.reshape(-1).copy()
    st, info = O.pspg_step(m, pos, q, dt=1e-3, solver_kind=2, rtol=1e-12)
    assert st == 0
    assert 1 <= info["picard_iters"] <= 4
    assert info["final_res"] <= 1e-6
    # fixed (wall) nodes did not move; fluid nodes moved by dt*v
    N = m.n_nodes
    wall = (m.flags & mg.FIXED) != 0
    np.testing.assert_array_equal(pos.reshape(2, N)[:, wall], m.pos[:, wall])
    np.testing.assert_allclose(pos.reshape(2, N)[:, ~wall], m.pos[:, ~wall] + 1e-3 * q.reshape(3, N)[:2, ~wall],
                               rtol=0, atol=1e-15)


def test_picard_failure_paths_restore():
    """maxIter = 0 with an unreachable minRes -> failure and restore (PicardAlgo.cpp:55-65)."""
    m = mg.dambreak_2d()
    q = mg.pspg_hydrostatic(m)
    q0 = q.copy()
    pos = m.pos.reshape(-1).copy()
    st, info = O.pspg_step(m, pos, q, dt=1e-3, min_res=0.0, max_iter=0, solver_kind=2, rtol=1e-12)
    assert st == 1
    np.testing.assert_array_equal(pos, m.pos.reshape(-1))
    np.testing.assert_array_equal(q, q0)


def unit_box_mesh(dim):
    """The unit square in 2 triangles or the unit cube in 6 Kuhn tetrahedra, each positively
    oriented (node order swapped where the determinant came out negative)."""
    if dim == 2:
        pos = np.array([[0.0, 1.0, 0.0, 1.0], [0.0, 0.0, 1.0, 1.0]])
        conn = [[0, 1, 3], [0, 3, 2]]
    else:
        pts = [(x, y, z) for z in (0, 1) for y in (0, 1) for x in (0, 1)]
        pos = np.array(pts, dtype=float).T
        conn = []
        for perm in itertools.permutations(range(3)):  # path 0 -> 7 along axes in this order
            v, path = 0, [0]
            for ax in perm:
                v |= 1 << ax
                path.append(v)
            conn.append(path)
    conn = np.array(conn, dtype=np.int32)
    for e in range(len(conn)):
        P = pos[:, conn[e]]
        if np.linalg.det(P[:, 1:] - P[:, :1]) < 0:
            conn[e, [0, 1]] = conn[e, [1, 0]]
    flags = np.zeros(pos.shape[1], np.uint8)
    return mg.PfemMesh(dim, pos, conn, flags, 1.0)


@pytest.mark.parametrize("dim", [2, 3])
def test_mass_closed_forms(dim):
    """Sum of element sizes = the box's measure; with a density linear in x the WC mass is the
    sum over elements of |e| times the mean nodal density, which for a linear field is the
    exact integral (the P1 mean is the centroid value): int_0^1 (2 + 3x) dx = 3.5."""
    m = unit_box_mesh(dim)
    assert O.mass(m, m.pos) == pytest.approx(1.0, abs=1e-14)
    rho = 2.0 + 3.0 * m.pos[0]
    assert O.mass(m, m.pos, rho) == pytest.approx(3.5, abs=1e-13)
    # uniform density scales the measure; a scaled box scales it by s^dim
    assert O.mass(m, m.pos, np.full(m.n_nodes, 1000.0)) == pytest.approx(1000.0, rel=1e-14)
    assert O.mass(m, 2.0 * m.pos) == pytest.approx(2.0 ** dim, rel=1e-14)


def test_minmax_matches_brute_force_and_keeps_the_zero_start():
    m = mg.dambreak_2d()
    fl = m.flags.copy()
    keep = (fl & (mg.BOUND | mg.FREE)) == 0
    assert keep.sum() > 0 and (~keep).sum() > 0
    for k in range(2):
        assert O.minmax(m, m.pos, k, "max") == max(0.0, m.pos[k][keep].max())
        assert O.minmax(m, m.pos, k, "min") == m.pos[k][keep].min()
    # the reference's largest value starts from 0: a domain below 0 reports 0
    shifted = m.pos - 10.0
    assert O.minmax(m, shifted, 1, "max") == 0.0
    assert O.minmax(m, shifted, 1, "min") == shifted[1][keep].min()


def moving_wall_mesh():
    """The 2D dam break with its left wall nodes un-fixed: their Dirichlet value depends on
    where they are, so it has to be re-evaluated at every Picard re-assembly."""
    m = mg.dambreak_2d(h=0.1)
    fl = m.flags.copy()
    left = ((fl & mg.BOUND) != 0) & (m.pos[0] < 0.5 * m.h)
    fl[left] &= ~np.uint8(mg.FIXED)
    m.flags = fl
    return m, left


def test_picard_position_dependent_dirichlet_fixed_point():
    """v_bc(x) = A x + c re-evaluated before every assembly: at convergence each moving
    Dirichlet node's velocity is the BC at the position that velocity moves it to, i.e.
    v = A (x0 + dt v) + c  ->  v = (I - dt A)^-1 (A x0 + c), while evaluating at the saved
    positions would give A x0 + c."""
    m, left = moving_wall_mesh()
    N = m.n_nodes
    A = np.array([[0.5, 0.4], [-0.3, 0.2]])
    c = np.array([0.01, -0.02])
    aff = np.hstack([A, c[:, None]])
    q = mg.pspg_hydrostatic(m)
    pos = m.pos.reshape(-1).copy()
    dt = 1e-3
    st, info = O.pspg_step(m, pos, q, dt=dt, residual="U", min_res=1e-12, max_iter=30, solver_kind=2, rtol=1e-13,
                           dir_affine=aff)
    assert st == 0, info
    x0 = m.pos[:, left]
    v = q.reshape(3, N)[:2, left]
    expect = np.linalg.solve(np.eye(2) - dt * A, A @ x0 + c[:, None])
    np.testing.assert_allclose(v, expect, rtol=0, atol=1e-9 * np.abs(expect).max())
    assert np.abs(expect - (A @ x0 + c[:, None])).max() > 1e-4 * np.abs(expect).max()
    # the moving wall nodes moved by dt v; the fixed ones stayed
    P = pos.reshape(2, N)
    np.testing.assert_allclose(P[:, left], x0 + dt * v, rtol=0, atol=1e-15)
    fixed = (m.flags & mg.FIXED) != 0
    np.testing.assert_array_equal(P[:, fixed], m.pos[:, fixed])


@pytest.mark.parametrize("dim", [2, 3])
def test_bingham_k_factor_closed_form(dim):
    """Bingham (MomContEquation.inl:100-123): only K changes, by mu_eq / mu; with an affine
    velocity the strain rate is uniform, gammaDot^2 = 2 sum eps_ii^2 + sum gamma_ij^2 (ddev of
    the incompressible solver), and mu_eq = tau0 (1 - exp(-m gammaDot)) / gammaDot; at rest
    mu_eq = tau0 m."""
    ne = dim + 1
    nd = dim * ne
    P = rand_simplex(dim, scale=0.05)
    A = np.array([[0.3, -1.2, 0.5], [0.8, -0.4, 0.2], [-0.6, 0.9, 0.1]])[:dim, :dim]
    vel = A @ P  # dim x ne
    qiter = np.concatenate([vel.reshape(-1), np.zeros(ne)])
    qprev = RNG.normal(size=(dim + 1) * ne)
    mu, tau0, m = 1e-3, 7.0, 300.0
    eps = 0.5 * (A + A.T)
    gd = math.sqrt(2 * np.sum(np.diag(eps) ** 2) + sum((A[i, j] + A[j, i]) ** 2 for i in range(dim)
                                                        for j in range(i + 1, dim)))
    for vel_scale, mu_eq in ((1.0, tau0 * (1 - math.exp(-m * gd)) / gd), (0.0, tau0 * m)):
        qi = qiter * vel_scale
        An, bn, taun = O.pspg_element(dim, P, qprev, qi, mu=mu)
        with O.material("Bingham", tau0=tau0, mReg=m):
            Ab, bb, taub = O.pspg_element(dim, P, qprev, qi, mu=mu)
        assert taub == taun  # tau keeps the material mu (MomContEquationPSPG.inl:257-258)
        np.testing.assert_array_equal(bb, bn)
        # non-momentum blocks unchanged; momentum block = M/dt + K(mu + mu_eq)
        np.testing.assert_array_equal(Ab[nd:], An[nd:])
        np.testing.assert_array_equal(Ab[:nd, nd:], An[:nd, nd:])
        vol = volume(P)
        Mdt = 1000.0 * vol * (np.ones((ne, ne)) + np.eye(ne)) / ((dim + 1) * (dim + 2)) / 1e-3
        Kn = An[:nd, :nd] - np.kron(np.eye(dim), Mdt)
        Kb = Ab[:nd, :nd] - np.kron(np.eye(dim), Mdt)
        np.testing.assert_allclose(Kb, Kn * (mu + mu_eq) / mu, rtol=1e-9, atol=1e-9 * np.abs(Kb).max())
