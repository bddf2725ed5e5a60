"""Fractional-step solver (Solver.id = "FracStep") on the GPU against the oracle
(oracle/pfem_oracle.cpp, fracStep): the closed-form accelerated linear flow, time steps of a
lid-driven cavity with free nodes, bound nodes without a Dirichlet value, gammaFS = 0.5 and a
Newtonian or Bingham fluid (velocities to 1e-9 of their scale, pressure to 1e-8 up to the
constant of the Laplacian's null vector, positions to 1e-12), the hydrostatic projection, and
the refusals of the C-ABI."""
import sys

import numpy as np
import pytest

import pfem_amd
from oracle import oracle as O
from pfem_amd import _lib
from pfem_amd import meshgen as mg
from test_oracle_frac_pins import FIXED, closed_box, gvec, lid_case

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def frac_case(ctx, dim, mesh, **kw):
    c = pfem_amd.MomContEqFracStep(ctx, dim, **kw)
    c.set_mesh(mesh)
    return c


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("res", ["U", "U_P"])
def test_frac_gpu_accelerated_linear_flow(ctx, dim, res):
    mesh, _ = closed_box(dim)
    mesh.flags = mesh.flags | FIXED
    P = mesh.pos.T
    N = mesh.n_nodes
    say(f"[frac] accelerated flow {dim}D {res} N={N}")
    rho, mu, dt = 1000.0, 0.1, 1e-2
    g = gvec(dim)
    acc = np.array([0.7, 0.3, -0.2][:dim])
    u = np.zeros((dim, N))
    u[0] = 0.3 * P[:, 1] + 0.05
    if dim == 3:
        u[1] = -0.2 * P[:, 2] + 0.02
    p_prev = rho * (P @ (g - acc))
    want_v = u + dt * acc[:, None]
    c = frac_case(ctx, dim, mesh, rho=rho, mu=mu, body_force=list(g), gammaFS=1.0, residual=res, krylov_rtol=1e-13)
    c.set_state(np.concatenate([u.ravel(), p_prev]))
    c.set_dirichlet(want_v)
    rc, st = c.step(dt)
    assert rc == _lib.PFEM_OK, st
    assert st["picard_iters"] >= 2
    q = c.get_state()
    v, p = q[: dim * N].reshape(dim, N), q[dim * N:]
    assert np.abs(v - want_v).max() < 1e-11
    assert np.ptp(p - p_prev) < 1e-10 * np.abs(p_prev).max()
    assert np.array_equal(c.get_positions(), mesh.pos)
    c.close()


@pytest.mark.parametrize("dim,bingham", [(2, None), (2, (5.0, 100.0)), (3, None), (3, (5.0, 100.0))])
def test_frac_gpu_lid_cavity_steps_match_oracle(ctx, dim, bingham):
    mesh, vbc = lid_case(dim)
    N = mesh.n_nodes
    say(f"[frac] lid cavity {dim}D bingham={bingham} N={N} E={mesh.n_elems}")
    rng = np.random.default_rng(7)
    rho, mu, dt, gam = 1000.0, 0.5, 2e-3, 0.5
    g = gvec(dim)
    v0 = 0.05 * rng.standard_normal((dim, N))
    dir_nodes = mesh.flags & mg.DIRICHLET > 0
    v0[:, dir_nodes] = vbc[:, dir_nodes]
    p0 = rho * (mesh.pos.T @ g) + 5 * rng.standard_normal(N)
    q_ref = np.concatenate([v0.ravel(), p0])
    pos_ref = mesh.pos.copy()
    mat = {} if bingham is None else dict(problem="Bingham", tau0=bingham[0], mReg=bingham[1])
    c = frac_case(ctx, dim, mesh, rho=rho, mu=mu, body_force=list(g), gammaFS=gam, residual="U", min_res=1e-8,
                  max_iter=20, krylov_rtol=1e-13, **mat)
    c.set_state(q_ref)
    c.set_dirichlet(vbc)
    used = np.zeros(N, bool)
    used[mesh.conn.ravel()] = True
    for step in range(3):
        if bingham is None:
            st_ref, s_ref = O.frac_step(mesh, pos_ref, q_ref, dir_vals=vbc.ravel().copy(), rho=rho, mu=mu, g=g, dt=dt,
                                        gamma_fs=gam, residual="U", min_res=1e-8, max_iter=20)
        else:
            with O.material("Bingham", tau0=bingham[0], mReg=bingham[1]):
                st_ref, s_ref = O.frac_step(mesh, pos_ref, q_ref, dir_vals=vbc.ravel().copy(), rho=rho, mu=mu, g=g,
                                            dt=dt, gamma_fs=gam, residual="U", min_res=1e-8, max_iter=20)
        rc, st = c.step(dt)
        say(f"  step {step}: oracle {st_ref} {s_ref['picard_iters']:.0f} it, gpu {rc} {st['picard_iters']} it "
            f"{st['krylov_iters']} cg")
        assert st_ref == 0 and rc == _lib.PFEM_OK
        assert st["picard_iters"] == s_ref["picard_iters"]
        q = c.get_state()
        v, p = q[: dim * N], q[dim * N:]
        v_ref, p_ref = q_ref[: dim * N], q_ref[dim * N:]
        assert np.abs(v - v_ref).max() < 1e-9 * np.abs(v_ref).max()
        dp = (p - p_ref)[used]
        assert np.ptp(dp) < 1e-8 * np.abs(p_ref).max()
        assert np.all(p[~used] == 0.0)
        assert np.abs(c.get_positions() - pos_ref).max() < 1e-12
    c.close()


@pytest.mark.parametrize("dim", [2, 3])
def test_frac_gpu_hydrostatic_projection_matches_oracle(ctx, dim):
    mesh, _ = closed_box(dim, n=15 if dim == 2 else 8)
    N = mesh.n_nodes
    say(f"[frac] hydrostatic projection {dim}D N={N}")
    rho, dt = 1000.0, 1e-2
    g = gvec(dim)
    q_ref = np.zeros((dim + 1) * N)
    pos_ref = mesh.pos.copy()
    st_ref, s_ref = O.frac_step(mesh, pos_ref, q_ref, rho=rho, mu=1e-3, g=g, dt=dt, gamma_fs=0.0, residual="U",
                                max_iter=3)
    c = frac_case(ctx, dim, mesh, rho=rho, mu=1e-3, body_force=list(g), gammaFS=0.0, residual="U", max_iter=3,
                  krylov_rtol=1e-13)
    rc, st = c.step(dt)
    assert st_ref == 0 and rc == _lib.PFEM_OK
    assert st["picard_iters"] == s_ref["picard_iters"]
    q = c.get_state()
    v, v_ref = q[: dim * N], q_ref[: dim * N]
    assert np.abs(v - v_ref).max() < 1e-9 * np.abs(v_ref).max()
    p, p_ref = q[dim * N:], q_ref[dim * N:]
    assert np.ptp(p - p_ref) < 1e-8 * np.abs(p_ref).max()
    c.close()


def test_frac_gpu_refusals(ctx):
    with pytest.raises(pfem_amd.PfemError) as e:
        pfem_amd.MomContEqFracStep(ctx, 2, residual="Ax_f")
    assert e.value.code == _lib.PFEM_E_UNSUPPORTED
    with pytest.raises(pfem_amd.PfemError):
        pfem_amd.MomContEqFracStep(ctx, 2, gammaFS=1.5)
    # Dirichlet nodes that move: their values would change with their positions every iteration
    mesh, _ = closed_box(2, fixed=False)
    c = frac_case(ctx, 2, mesh)
    rc = _lib.lib().pfem_frac_step(c._h, 1e-3, None)
    assert rc == _lib.PFEM_E_UNSUPPORTED
    # the batched and serving paths step PSPG and WC cases only
    mesh, _ = closed_box(2)
    c2 = frac_case(ctx, 2, mesh)
    with pytest.raises(pfem_amd.PfemError) as e:
        pfem_amd.batch_step(ctx, [c2], [1e-3])
    assert e.value.code == _lib.PFEM_E_UNSUPPORTED
    c.close()
    c2.close()
