"""Pins of the oracle's weakly compressible step, its CFL time step and the Picard residuals.

Each check is a closed form or an invariant that the mathematics of the P1 discretisation
fixes independently of how the oracle computes it, chosen so that a plausible restatement
slip fails it:

* continuity (ContEquation.inl:353-413): with a uniform pressure and an affine velocity,
  every node's new pressure is p0 - dt (K0 + K0' p0) div v (a K0/K0' swap, a sign slip or a
  dropped D factor fails); with v = 0 and a linear pressure, Meduri stabilisation gives the
  consistent-mass average and None leaves p unchanged (a Meduri/lumped swap fails), and the
  lumped-mass-weighted pressure is conserved;
* lumped masses (MatricesBuilder.hpp:87-129): sum lump2(M) is the domain's measure and
  sum lump(M(rho)) is the exact integral of a linear density;
* momentum (MomEquation.inl:229-302): a uniform strain rate gives zero force at interior
  nodes and the traction -h^(d-1) sigma n at face nodes with sigma = mu ddev eps, which pins
  ddev (MomEquation.inl:63-83) and the Voigt order of B; a hydrostatic column is at rest
  (interior a = 0: a flipped D^T P, a missing or mis-scaled F(rho) fails); the body force
  sums to g times the mass;
* the Tait EOS (ContEquation.inl:319-332) through a full step, on exactly known answers;
* computeNextDT (WCompNewton/Solver.cpp:192-234) where u^2, the viscous term 4 alpha^2/h^2
  or the K0' p part of c^2 dominates;
* the U and U_P residuals (MomContEquationPSPG.inl:314-360) on hand-built iterates with
  hand-computed values.
"""
import math

import numpy as np
import pytest

from oracle import oracle as O
from pfem_amd import meshgen as mg


def box_mesh(dim, h=0.1, size=None, jitter=True, seed=3):
    """Lattice box without walls: every node moves and belongs to an element."""
    if dim == 2:
        m = mg.tank_2d(h=h, size=size or (0.6, 0.4), fill=1.0, seed=seed)
    else:
        m = mg.tank_3d(h=h, size=size or (0.4, 0.3, 0.3), fill=1.0, seed=seed)
    if not jitter:
        m.pos = np.round(m.pos / h) * h
    m.flags[:] = 0
    return m


def element_volumes(m, pos=None):
    P = m.pos if pos is None else pos
    X = P[:, m.conn]  # dim x E x (dim+1)
    J = X[:, :, 1:] - X[:, :, :1]  # dim x E x dim
    return np.linalg.det(np.moveaxis(J, 1, 0)) / math.factorial(m.dim)


def face_class(m, h, size):
    """For each node: -1 interior, else the index of the single box face it lies on
    (2*axis + (0 low / 1 high)); nodes on two or more faces are -2."""
    lat = np.round(m.pos / h).astype(int)
    hi = [int(round(s / h)) for s in size]
    out = np.full(m.n_nodes, -1)
    for n in range(m.n_nodes):
        faces = [2 * k + (1 if lat[k, n] == hi[k] else 0) for k in range(m.dim) if lat[k, n] in (0, hi[k])]
        out[n] = -1 if not faces else (faces[0] if len(faces) == 1 else -2)
    return out


# ------------------------------------------------------------------------------------------
# continuity
# ------------------------------------------------------------------------------------------
@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("meduri", [True, False])
def test_wc_continuity_uniform_pressure_affine_velocity(dim, meduri):
    m = box_mesh(dim)
    N = m.n_nodes
    rng = np.random.default_rng(7)
    A = rng.normal(size=(dim, dim)) * 0.5
    c = rng.normal(size=dim) * 0.1
    v = (A @ m.pos + c[:, None]).reshape(-1)
    K0, K0p, p0, dt = 2.2e5, 7.6, 3.0e4, 1e-3
    p = np.full(N, p0)
    rho = np.full(N, 1000.0)
    a = np.zeros(dim * N)
    pos = m.pos.reshape(-1).copy()
    O.wc_step(m, pos, v, p, rho, a, dt, g=[0.0] * dim, K0=K0, K0p=K0p, rho_star=1000.0, meduri=meduri)
    # positions moved by dt v (a = 0): x' = (I + dt A) x + dt c, so on the moved mesh the
    # velocity is affine with gradient A (I + dt A)^-1
    div = np.trace(A @ np.linalg.inv(np.eye(dim) + dt * A))
    expect = p0 - dt * (K0 + K0p * p0) * div
    scale = dt * (K0 + K0p * p0) * abs(div)
    assert scale > 1e-3 * p0
    np.testing.assert_allclose(p, expect, rtol=0, atol=1e-10 * scale + 1e-12 * p0)


@pytest.mark.parametrize("dim", [2, 3])
def test_wc_continuity_linear_pressure_meduri_vs_none(dim):
    """v = 0: None keeps p; Meduri replaces it by the consistent-mass average
    sum_e |e| (p_i + sum_j p_j) / ((d+1)(d+2)) / sum_e |e|/(d+1)."""
    m = box_mesh(dim)
    N = m.n_nodes
    vol = element_volumes(m)
    p0 = 1e4 + 3e3 * m.pos[0] - 2e3 * m.pos[dim - 1] + 500.0 * np.sin(7 * m.pos[0])
    ml = np.zeros(N)
    num = np.zeros(N)
    ne = dim + 1
    for e in range(m.n_elems):
        nodes = m.conn[e]
        s = p0[nodes].sum()
        for i in nodes:
            ml[i] += vol[e] / ne
            num[i] += vol[e] * (p0[i] + s) / (ne * (ne + 1))
    for meduri, expect in ((False, p0), (True, num / ml)):
        p = p0.copy()
        pos = m.pos.reshape(-1).copy()
        O.wc_step(m, pos, np.zeros(dim * N), p, np.full(N, 1000.0), np.zeros(dim * N), 1e-4, g=[0.0] * dim,
                  K0=2.2e9, K0p=7.0, rho_star=1000.0, meduri=meduri)
        np.testing.assert_allclose(p, expect, rtol=1e-13, atol=0)
        if meduri:  # sum ML p is conserved (M symmetric)
            assert float(ml @ p) == pytest.approx(float(ml @ p0), rel=1e-13)
    # the two stabilisations differ on this field
    assert np.abs(num / ml - p0).max() > 1.0


@pytest.mark.parametrize("dim", [2, 3])
def test_wc_tait_eos_known_answers_through_a_step(dim):
    """Uniform pressure, v = 0: p stays p0 and rho = rho* ((K0'/K0) p0 + 1)^(1/K0'):
    K0' = 1, p0 = K0 gives 2 rho*; K0' = 2, p0 = 4 K0 gives 3 rho*."""
    m = box_mesh(dim)
    N = m.n_nodes
    for K0, K0p, p0, ratio in ((2.2e5, 1.0, 2.2e5, 2.0), (1.5e6, 2.0, 6.0e6, 3.0)):
        p = np.full(N, p0)
        rho = np.full(N, 1.0)
        pos = m.pos.reshape(-1).copy()
        O.wc_step(m, pos, np.zeros(dim * N), p, rho, np.zeros(dim * N), 1e-4, g=[0.0] * dim, K0=K0, K0p=K0p,
                  rho_star=998.0)
        np.testing.assert_allclose(p, p0, rtol=1e-13)
        np.testing.assert_allclose(rho, ratio * 998.0, rtol=1e-13)


# ------------------------------------------------------------------------------------------
# lumped masses and forces
# ------------------------------------------------------------------------------------------
@pytest.mark.parametrize("dim", [2, 3])
def test_wc_lumped_mass_sums(dim):
    size = (0.6, 0.4) if dim == 2 else (0.4, 0.3, 0.3)
    m = box_mesh(dim, jitter=False)
    N = m.n_nodes
    measure = float(np.prod(size))
    coef = np.array([1000.0, 50.0, 20.0, -30.0][:dim + 1])
    rho = coef[0] + coef[1:] @ m.pos
    # exact integral of a linear density over the box: measure * rho(centre)
    rho_int = measure * (coef[0] + coef[1:] @ (np.array(size) / 2))
    g = np.array([0.3, -9.81, 1.7][:dim])
    ML, F0, Mv, F = O.wc_parts(m, m.pos, np.zeros(dim * N), np.zeros(N), rho, g=list(g), mu=1e-3)
    assert ML.sum() == pytest.approx(measure, rel=1e-13)
    for k in range(dim):
        assert Mv[k * N:(k + 1) * N].sum() == pytest.approx(rho_int, rel=1e-13)
        # v = 0, p = 0: the force is the body force alone, summing to g_k times the mass
        assert F[k * N:(k + 1) * N].sum() == pytest.approx(g[k] * rho_int, rel=1e-12)
    np.testing.assert_array_equal(F0, 0.0)
    # uniform density: each node's lumped momentum mass is rho times its lumped volume
    ML1, _, Mv1, _ = O.wc_parts(m, m.pos, np.zeros(dim * N), np.zeros(N), np.full(N, 1000.0), g=list(g))
    for k in range(dim):
        np.testing.assert_allclose(Mv1[k * N:(k + 1) * N], 1000.0 * ML1, rtol=1e-13)


@pytest.mark.parametrize("dim", [2, 3])
def test_wc_uniform_strain_rate_traction(dim):
    """v = A x, p = 0, g = 0: sigma = mu ddev eps is uniform, so F = -K V vanishes at interior
    nodes and equals -h^(d-1) sigma n at nodes inside a box face (the hat function integrates
    to h^(d-1) over the face)."""
    h = 0.1
    size = (0.6, 0.4) if dim == 2 else (0.4, 0.3, 0.3)
    m = box_mesh(dim, h=h, jitter=False)
    N = m.n_nodes
    mu = 0.37
    A = np.array([[0.8, 0.3, -0.4], [-0.9, -0.2, 0.6], [0.5, 0.7, 0.35]])[:dim, :dim]
    v = (A @ m.pos).reshape(-1)
    _, _, _, F = O.wc_parts(m, m.pos, v, np.zeros(N), np.full(N, 1000.0), g=[0.0] * dim, mu=mu)
    F = F.reshape(dim, N)
    eps = 0.5 * (A + A.T)
    # WC ddev: 4/3, -2/3 on the normal block (trace convention of 3D also in 2D), 1 on the
    # engineering shears -> sigma = mu (2 eps - 2/3 tr(eps) I)
    sigma = mu * (2 * eps - 2.0 / 3.0 * np.trace(eps) * np.eye(dim))
    fc = face_class(m, h, size)
    scale = mu * np.abs(A).max() * h ** (dim - 1)
    interior = fc == -1
    assert interior.sum() > 0
    np.testing.assert_allclose(F[:, interior], 0.0, atol=1e-12 * scale)
    nface = 0
    for face in range(2 * dim):
        sel = fc == face
        nface += sel.sum()
        normal = np.zeros(dim)
        normal[face // 2] = 1.0 if face % 2 else -1.0
        expect = -h ** (dim - 1) * sigma @ normal
        np.testing.assert_allclose(F[:, sel], np.repeat(expect[:, None], sel.sum(), axis=1), rtol=0,
                                   atol=1e-12 * scale)
    assert nface > 0
    # a rigid rotation has no strain rate: no force anywhere
    W = np.zeros((dim, dim))
    W[0, 1], W[1, 0] = 0.7, -0.7
    _, _, _, Fr = O.wc_parts(m, m.pos, (W @ m.pos).reshape(-1), np.zeros(N), np.full(N, 1000.0), g=[0.0] * dim,
                             mu=mu)
    np.testing.assert_allclose(Fr, 0.0, atol=1e-12 * scale)


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("jitter", [False, True])
def test_wc_hydrostatic_column_at_rest(dim, jitter):
    """v = 0, p = rho* |g| (H - z), K0 so large that rho = rho* to ~1e-11: D^T P balances
    F(rho) at every node not on a wall (also on the free surface, where p = 0 on the
    lattice), so a = 0 there; on the un-jittered lattice the bottom nodes' acceleration is
    the closed form of the pressure on their face."""
    h = 0.1
    m = box_mesh(dim, h=h, jitter=jitter)
    N = m.n_nodes
    H = float(np.round(m.pos[dim - 1] / h).max() * h)
    g = 9.81
    rs = 1000.0
    p = rs * g * (H - m.pos[dim - 1])
    rho = np.full(N, rs)
    v = np.zeros(dim * N)
    a = np.zeros(dim * N)
    pos = m.pos.reshape(-1).copy()
    gv = [0.0] * dim
    gv[-1] = -g
    O.wc_step(m, pos, v, p, rho, a, 1e-5, g=gv, mu=1e-3, K0=1e15, K0p=7.0, rho_star=rs, meduri=False)
    a = a.reshape(dim, N)
    lat = np.round(m.pos / h).astype(int)
    hi = np.round(np.array((0.6, 0.4) if dim == 2 else (0.4, 0.3, 0.3)) / h).astype(int)
    on_wall = np.zeros(N, bool)
    for k in range(dim - 1):
        on_wall |= (lat[k] == 0) | (lat[k] == hi[k])
    on_wall |= lat[dim - 1] == 0
    top = lat[dim - 1] == hi[dim - 1]
    check = ~on_wall if not jitter else (~on_wall & ~top)
    assert check.sum() > 0
    np.testing.assert_allclose(a[:, check], 0.0, atol=1e-8 * g)
    if not jitter:
        # an unconstrained bottom node feels the pressure on its face, -p(0) h^(d-1) n, over a
        # lumped mass rho h^d / 2 (and the interior balance): a_z = -2 g H / h, a_x = 0
        bottom = (lat[dim - 1] == 0) & ~np.any([(lat[k] == 0) | (lat[k] == hi[k]) for k in range(dim - 1)],
                                               axis=0)
        assert bottom.sum() > 0
        np.testing.assert_allclose(a[dim - 1, bottom], -2 * g * H / h, rtol=1e-9)
        np.testing.assert_allclose(a[:dim - 1, bottom], 0.0, atol=1e-8 * g)


# ------------------------------------------------------------------------------------------
# computeNextDT
# ------------------------------------------------------------------------------------------
def tet_inradius(P):
    """3 V / (total face area) of a tetrahedron given as 3 x 4 corner coordinates."""
    V = abs(np.linalg.det(P[:, 1:] - P[:, :1])) / 6
    area = 0.0
    for f in ((0, 1, 2), (0, 1, 3), (0, 2, 3), (1, 2, 3)):
        area += 0.5 * np.linalg.norm(np.cross(P[:, f[1]] - P[:, f[0]], P[:, f[2]] - P[:, f[0]]))
    return 3 * V / area


@pytest.mark.parametrize("dim", [2, 3])
def test_wc_cfl_branches(dim):
    """dt = min_e sc he / sqrt(max_n max(u^2, c^2, 4 alpha^2 / he^2)) on an un-jittered lattice
    (one element shape, he = 2 r_in) in each of its branches."""
    h = 0.1
    m = box_mesh(dim, h=h, jitter=False)
    N = m.n_nodes
    if dim == 2:
        he = h * (2 - math.sqrt(2))
    else:
        he = 2 * min(tet_inradius(m.pos[:, m.conn[e]]) for e in range(m.n_elems))
        assert 2 * max(tet_inradius(m.pos[:, m.conn[e]]) for e in range(m.n_elems)) == pytest.approx(he, rel=1e-12)
    sc = 0.5
    pos = m.pos.reshape(-1).copy()
    rho = np.full(N, 1000.0)
    zero_v = np.zeros(dim * N)
    # u^2 dominates: uniform velocity (3, 4[, 12]) -> |u| = 5 (13)
    u = np.array([3.0, 4.0, 12.0][:dim])
    v = np.repeat(u[:, None], N, axis=1).reshape(-1)
    dt = O.wc_next_dt(m, pos, v, np.zeros(N), rho, mu=0.0, K0=1e-9, K0p=1e-9, security_coeff=sc, max_dt=1.0)
    assert dt == pytest.approx(sc * he / np.linalg.norm(u), rel=1e-12)
    # the largest speed of any node wins; it sets dt through every element holding it
    v2 = np.zeros((dim, N))
    v2[0, N // 2] = 20.0
    dt = O.wc_next_dt(m, pos, v2.reshape(-1), np.zeros(N), rho, mu=0.0, K0=1e-9, K0p=1e-9, security_coeff=sc,
                      max_dt=1.0)
    assert dt == pytest.approx(sc * he / 20.0, rel=1e-12)
    # viscous term: alpha = mu / rho, dt = sc he^2 / (2 alpha)
    mu = 10.0
    dt = O.wc_next_dt(m, pos, zero_v, np.zeros(N), rho, mu=mu, K0=1e-9, K0p=1e-9, security_coeff=sc, max_dt=1.0)
    assert dt == pytest.approx(sc * he * he / (2 * mu / 1000.0), rel=1e-12)
    # c^2 = (K0 + K0' p) / rho with p dominant in it
    K0, K0p, p0 = 2.2e5, 7.6, 1.0e5
    dt = O.wc_next_dt(m, pos, zero_v, np.full(N, p0), rho, mu=0.0, K0=K0, K0p=K0p, security_coeff=sc, max_dt=1.0)
    assert dt == pytest.approx(sc * he / math.sqrt((K0 + K0p * p0) / 1000.0), rel=1e-12)
    # maxDT caps it
    assert O.wc_next_dt(m, pos, zero_v, np.zeros(N), rho, mu=0.0, K0=K0, K0p=K0p, security_coeff=sc,
                        max_dt=1e-9) == 1e-9


# ------------------------------------------------------------------------------------------
# Picard residuals
# ------------------------------------------------------------------------------------------
def test_picard_residual_u_and_u_p_hand_values():
    """3 nodes in 2D, node 2 free (excluded).  Field-blocked q = (u, v, p):
    prev velocities (3, 4), (0, 0); iterate (4, 6), (2, -1): num = 1+4+4+1 = 10, den = 25
    -> U = sqrt(0.4).  Pressures prev (2, 0), iterate (2, 3): num 9, den 4 -> 1.5."""
    flags = np.array([0, mg.BOUND, mg.FREE], np.uint8)
    #                 u0   u1   u2     v0   v1    v2     p0   p1   p2
    prev = np.array([3.0, 0.0, 100.0, 4.0, 0.0, -50.0, 2.0, 0.0, 7.0])
    it = np.array([4.0, 2.0, -80.0, 6.0, -1.0, 20.0, 2.0, 3.0, 99.0])
    assert O.picard_residual(2, flags, "U", it, prev) == pytest.approx(math.sqrt(0.4), rel=1e-15)
    assert O.picard_residual(2, flags, "U_P", it, prev) == pytest.approx(1.5, rel=1e-15)
    # U_P with the velocity ratio the larger one
    it2 = it.copy()
    it2[7] = 0.5  # pressure num 0.25, den 4 -> 0.25 < sqrt(0.4)
    assert O.picard_residual(2, flags, "U_P", it2, prev) == pytest.approx(math.sqrt(0.4), rel=1e-15)
    # zero denominators give DBL_MAX (the first Picard iteration from a state at rest)
    big = np.finfo(np.float64).max
    assert O.picard_residual(2, flags, "U", it, np.zeros(9)) == big
    z = prev.copy()
    z[6:8] = 0.0
    assert O.picard_residual(2, flags, "U_P", it, z) == big


def test_picard_residual_axf_is_the_true_residual_norm():
    m = mg.dambreak_2d(h=0.2)
    N = m.n_nodes
    q = mg.pspg_hydrostatic(m)
    S = O.pspg_assemble(m, q, dt=1e-3)
    rng = np.random.default_rng(5)
    x = rng.normal(size=S.n)
    A = S.to_scipy()
    r = O.picard_residual(2, m.flags, "Ax_f", x, np.zeros(S.n), system=S)
    assert r == pytest.approx(float(np.linalg.norm(A @ x - S.b)), rel=1e-13)
    assert N * 3 == S.n
