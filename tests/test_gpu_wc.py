"""GPU parity of the weakly compressible CDS_dpdt step and the CFL time step against the
CPU oracle (FP64 state: 1e-10 normwise per field; FP32 state: 2e-5)."""
import dataclasses
import sys

import numpy as np
import pytest

import pfem_amd
from oracle import oracle as O
from pfem_amd import meshgen as mg

pytestmark = pytest.mark.gpu

K0, K0P, RS, MU = 2.2e6, 7.0, 1000.0, 1e-3


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def meshes():
    return {
        "2d_tank": mg.tank_2d(h=0.05, size=(1.0, 0.5), fill=0.6, seed=2),
        "3d_tank": mg.tank_3d(h=0.05, size=(0.4, 0.4, 0.3), fill=0.7, seed=3),
    }


def rel(ref, got):
    s = np.abs(ref).max()
    return np.abs(got - ref).max() / (s if s > 0 else 1.0)


@pytest.mark.parametrize("state", ["fp64", "fp32"])
@pytest.mark.parametrize("name", ["2d_tank", "3d_tank"])
def test_wc_step_parity(ctx, name, state):
    mesh = meshes()[name]
    say(f"[wc] {name} {state} N={mesh.n_nodes} E={mesh.n_elems}")
    d, N = mesh.dim, mesh.n_nodes
    v, p, rho, a = mg.wc_tait_hydrostatic(mesh, K0=K0, K0p=K0P, rho_star=RS)
    c = pfem_amd.Case(ctx, d, kind="WC", mu=MU, K0=K0, K0p=K0P, rho_star=RS, wc_state=state)
    c.set_mesh(mesh)
    c.set_state(np.concatenate([v.reshape(-1), p, rho, a.reshape(-1)]))
    pos_o = mesh.pos.reshape(-1).copy()
    vo, po, ro, ao = v.reshape(-1).copy(), p.copy(), rho.copy(), a.reshape(-1).copy()
    tol = 1e-10 if state == "fp64" else 2e-5
    for step in range(3):
        dt_o = O.wc_next_dt(mesh, pos_o, vo, po, ro, mu=MU, K0=K0, K0p=K0P, security_coeff=0.5, max_dt=1e-3)
        dt_g = c.wc_next_dt(0.5, 1e-3)
        assert dt_g == pytest.approx(dt_o, rel=1e-12 if state == "fp64" else 1e-5)
        O.wc_step(mesh, pos_o, vo, po, ro, ao, dt_o, mu=MU, K0=K0, K0p=K0P, rho_star=RS)
        c.wc_step(dt_o)
        s = c.get_state()
        vg, pg, rg, ag = s[:d * N], s[d * N:(d + 1) * N], s[(d + 1) * N:(d + 2) * N], s[(d + 2) * N:]
        say(f"  step {step}: dv={rel(vo, vg):.2e} dp={rel(po, pg):.2e} drho={rel(ro, rg):.2e} da={rel(ao, ag):.2e}")
        assert rel(vo, vg) <= tol
        assert rel(po, pg) <= tol
        assert rel(ro, rg) <= tol
        assert np.abs(ag - ao).max() <= tol * max(np.abs(ao).max(), 9.81)
        # positions are FP64 in both states; with FP32 velocities they drift by dt * |dv|
        np.testing.assert_allclose(c.get_positions().reshape(-1), pos_o, rtol=0,
                                   atol=1e-12 if state == "fp64" else 1e-9)


@pytest.mark.parametrize("state", ["fp64", "fp32"])
@pytest.mark.parametrize("name", ["2d_tank", "3d_tank"])
def test_wc_step_parity_dirichlet_without_stabilization(ctx, name, state):
    """Wall nodes on the x = min side flagged Dirichlet with random values (the reference's
    'a = v_bc' rule of MomEquation.inl:358-364) and ContEq.stabilization = "None" (lumped
    continuity mass instead of Meduri's consistent one), against the oracle."""
    base = meshes()[name]
    d, N = base.dim, base.n_nodes
    flags = base.flags.copy()
    left = ((flags & mg.BOUND) != 0) & (base.pos[0] < base.pos[0].min() + 0.5 * base.h)
    flags[left] |= mg.DIRICHLET
    mesh = dataclasses.replace(base, flags=flags)
    say(f"[wc-dir] {name} {state} N={N} dirichlet={int(left.sum())}")
    assert left.sum() > 0
    dv = np.random.default_rng(7).normal(scale=0.05, size=d * N)
    v, p, rho, a = mg.wc_tait_hydrostatic(mesh, K0=K0, K0p=K0P, rho_star=RS)
    c = pfem_amd.Case(ctx, d, kind="WC", mu=MU, K0=K0, K0p=K0P, rho_star=RS, wc_state=state,
                      stabilization="None")
    c.set_mesh(mesh)
    c.set_state(np.concatenate([v.reshape(-1), p, rho, a.reshape(-1)]))
    c.set_dirichlet(dv)
    pos_o = mesh.pos.reshape(-1).copy()
    vo, po, ro, ao = v.reshape(-1).copy(), p.copy(), rho.copy(), a.reshape(-1).copy()
    tol = 1e-10 if state == "fp64" else 2e-5
    for step in range(3):
        dt = 2e-5
        O.wc_step(mesh, pos_o, vo, po, ro, ao, dt, dir_vals=dv, mu=MU, K0=K0, K0p=K0P, rho_star=RS, meduri=False)
        c.wc_step(dt)
        s = c.get_state()
        vg, pg, rg, ag = s[:d * N], s[d * N:(d + 1) * N], s[(d + 1) * N:(d + 2) * N], s[(d + 2) * N:]
        say(f"  step {step}: dv={rel(vo, vg):.2e} dp={rel(po, pg):.2e} drho={rel(ro, rg):.2e} da={rel(ao, ag):.2e}")
        assert rel(vo, vg) <= tol
        assert rel(po, pg) <= tol
        assert rel(ro, rg) <= tol
        assert np.abs(ag - ao).max() <= tol * max(np.abs(ao).max(), 9.81)
        # Dirichlet nodes: a holds the prescribed values (as the reference's solve leaves it)
        for k in range(d):
            np.testing.assert_allclose(ag[k * N:(k + 1) * N][left], dv[k * N:(k + 1) * N][left],
                                       rtol=0, atol=1e-12 if state == "fp64" else 1e-7)


@pytest.mark.parametrize("state", ["fp64", "fp32"])
@pytest.mark.parametrize("name", ["2d_tank", "3d_tank"])
def test_wc_mass_matches_oracle(ctx, name, state):
    """The WC global mass (element sizes times mean nodal density) on the device against the
    oracle, before and after steps (FP32 state: the oracle reads the same FP32-rounded rho)."""
    mesh = meshes()[name]
    d, N = mesh.dim, mesh.n_nodes
    say(f"[wc-mass] {name} {state}")
    v, p, rho, a = mg.wc_tait_hydrostatic(mesh, K0=K0, K0p=K0P, rho_star=RS)
    c = pfem_amd.Case(ctx, d, kind="WC", mu=MU, K0=K0, K0p=K0P, rho_star=RS, wc_state=state)
    c.set_mesh(mesh)
    c.set_state(np.concatenate([v.reshape(-1), p, rho, a.reshape(-1)]))
    for step in range(3):
        s = c.get_state()
        rg = s[(d + 1) * N:(d + 2) * N]
        m_ref = O.mass(mesh, c.get_positions().reshape(d, -1), rg)
        m_gpu = c.mass()
        say(f"  step {step}: mass {m_gpu:.15g} vs {m_ref:.15g}")
        assert m_gpu == pytest.approx(m_ref, rel=1e-12)
        c.wc_step(c.wc_next_dt(0.5, 1e-3))


def test_wc_full_size_matches_oracle(ctx):
    """The C3 sloshing tank at its full size (2x2x1 m half filled, h = 0.01 m: 2.06M nodes,
    12M tets), FP64 state: two CDS_dpdt steps with computeNextDT against the oracle."""
    mesh = mg.tank_3d(h=0.01)
    d, N = mesh.dim, mesh.n_nodes
    say(f"[wc-full] N={N} E={mesh.n_elems}")
    assert N > 2_000_000 and mesh.n_elems == 12_000_000
    v, p, rho, a = mg.wc_tait_hydrostatic(mesh, K0=2.2e9, K0p=7.0, rho_star=RS)
    c = pfem_amd.Case(ctx, d, kind="WC", mu=MU, K0=2.2e9, K0p=7.0, rho_star=RS, wc_state="fp64")
    c.set_mesh(mesh)
    c.set_state(np.concatenate([v.reshape(-1), p, rho, a.reshape(-1)]))
    pos_o = mesh.pos.reshape(-1).copy()
    vo, po, ro, ao = v.reshape(-1).copy(), p.copy(), rho.copy(), a.reshape(-1).copy()
    for step in range(2):
        dt_o = O.wc_next_dt(mesh, pos_o, vo, po, ro, mu=MU, K0=2.2e9, K0p=7.0, security_coeff=0.5, max_dt=1e-3)
        dt_g = c.wc_next_dt(0.5, 1e-3)
        assert dt_g == pytest.approx(dt_o, rel=1e-12)
        O.wc_step(mesh, pos_o, vo, po, ro, ao, dt_o, mu=MU, K0=2.2e9, K0p=7.0, rho_star=RS)
        c.wc_step(dt_o)
        s = c.get_state()
        vg, pg, rg, ag = s[:d * N], s[d * N:(d + 1) * N], s[(d + 1) * N:(d + 2) * N], s[(d + 2) * N:]
        say(f"  step {step}: dt={dt_g:.6e} dv={rel(vo, vg):.2e} dp={rel(po, pg):.2e} drho={rel(ro, rg):.2e} "
            f"da={rel(ao, ag):.2e}")
        assert rel(vo, vg) <= 1e-10
        assert rel(po, pg) <= 1e-10
        assert rel(ro, rg) <= 1e-10
        assert np.abs(ag - ao).max() <= 1e-10 * max(np.abs(ao).max(), 9.81)
        np.testing.assert_allclose(c.get_positions().reshape(-1), pos_o, rtol=0, atol=1e-12)
    assert c.mass() == pytest.approx(O.mass(mesh, pos_o.reshape(d, -1), ro), rel=1e-12)
    c.close()
