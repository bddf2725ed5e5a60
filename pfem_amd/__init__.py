"""pfem_amd — MI355X (gfx950) PFEM per-time-step assembly and sparse solve.

Python host mirror of the C-ABI in include/pfem_gpu.h.  The classes follow the
reference's equation/solver interface for this path (ImperatorS79/PFEM):

* ``MomContEqIncompNewton`` — the PSPG momentum-continuity equation whose ``solve()``
  runs one Picard time step (``MomContEquation.inl:287-298``) and returns ``False`` on
  numerical failure, raising on configuration errors like the reference.
* ``WCompNewtonStep`` — the weakly compressible CDS_dpdt step
  (``WCompNewton/Solver.cpp:236-276``) and ``computeNextDT`` (``:192-234``).
* ``Case`` exposes the building blocks (assemble, solve, CSR export) used by the parity
  tests, ``batch_step`` runs many independent cases as one batch, and ``Server`` is the
  continuous-batching request loop (one request = one time step of one case).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import PfemError, check, lib
from .meshgen import BOUND, DIRICHLET, FIXED, FREE, ONFS, PfemMesh

__all__ = ["Context", "Case", "MomContEqIncompNewton", "MomContEqFracStep", "PicardAlgo", "WCompNewtonStep",
           "HeatEqIncompNewton", "BoussinesqIncompNewton", "SolverIncompNewton", "Mesher", "batch_step", "Server",
           "now_ns",
           "PfemError", "PfemMesh", "BOUND", "FREE", "ONFS", "FIXED", "DIRICHLET"]


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _ip(a):
    return a.ctypes.data_as(C.POINTER(C.c_int32))


def _u8(a):
    return a.ctypes.data_as(C.POINTER(C.c_uint8))


class Context:
    """One per process per GPU (pfem_ctx_create)."""

    def __init__(self, device: int = 0):
        self._h = C.c_void_p()
        check(lib().pfem_ctx_create(device, C.byref(self._h)))
        self.device = device

    def xcd_map(self, nblocks: int = 2048) -> np.ndarray:
        """XCD id each of nblocks one-wavefront workgroups ran on (pfem_probe_xcd): the
        block-to-XCD mapping the XCD-aware kernel segments assume (workgroup g on XCD g mod 8)."""
        out = np.empty(nblocks, dtype=np.int32)
        check(lib().pfem_probe_xcd(self._h, int(nblocks), _ip(out)))
        return out

    def close(self):
        if self._h:
            lib().pfem_ctx_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Case:
    """One independent PFEM case resident on the device (pfem_case_create)."""

    def __init__(self, ctx: Context, dim: int, kind: str = "PSPG", rho=1000.0, mu=1e-3, gamma=0.0, body_force=None,
                 K0=2.2e9, K0p=7.0, rho_star=1000.0, stabilization="Meduri", wc_state="fp64", residual="Ax_f",
                 max_iter=10, min_res=1e-6, precision="fp64", krylov_rtol=1e-9, krylov_max_iter=20000,
                 inner_rtol=1e-3, inner_max_iter=300, precond="auto", problem="Newtonian", tau0=0.0, mReg=0.0,
                 cv=0.0, k=0.0, h=0.0, Tinf=0.0, epsRad=0.0, gammaFS=0.0, alpha=0.0, Tr=0.0, DgammaDT=0.0):
        d = _lib.CaseDesc()
        d.dim = dim
        kinds = {"PSPG": _lib.KIND_PSPG, "WC": _lib.KIND_WC, "HEAT": _lib.KIND_HEAT, "FracStep": _lib.KIND_FRACSTEP,
                 "Boussinesq": _lib.KIND_BOUSSINESQ}
        if kind not in kinds:
            raise PfemError(_lib.PFEM_E_INVALID, "unknown case kind: " + str(kind))
        d.kind = kinds[kind]
        d.cv, d.k_cond, d.h_conv, d.t_inf, d.eps_rad = cv, k, h, Tinf, epsRad
        d.rho, d.mu, d.gamma = rho, mu, gamma
        d.gamma_fs = gammaFS
        d.alpha_exp, d.t_ref, d.dgamma_dt = alpha, Tr, DgammaDT
        bf = body_force if body_force is not None else ([0.0, -9.81] if dim == 2 else [0.0, 0.0, -9.81])
        for k in range(3):
            d.body_force[k] = bf[k] if k < len(bf) else 0.0
        d.K0, d.K0p, d.rho_star = K0, K0p, rho_star
        if stabilization not in ("Meduri", "None"):
            raise PfemError(_lib.PFEM_E_INVALID, "unknown stabilization: " + str(stabilization))
        d.meduri = 1 if stabilization == "Meduri" else 0
        d.wc_state = _lib.STATE_FP32 if wc_state == "fp32" else _lib.STATE_FP64
        if residual not in _lib.RES:
            raise PfemError(_lib.PFEM_E_INVALID, "unknown residual type: " + str(residual))
        d.residual = _lib.RES[residual]
        d.max_iter, d.min_res = max_iter, min_res
        if precision not in _lib.PRECISION:
            raise PfemError(_lib.PFEM_E_INVALID, "unknown precision: " + str(precision))
        d.precision = _lib.PRECISION[precision]
        d.krylov_max_iter, d.krylov_rtol, d.inner_rtol = krylov_max_iter, krylov_rtol, inner_rtol
        d.inner_max_iter = inner_max_iter
        if precond not in _lib.PRECOND:
            raise PfemError(_lib.PFEM_E_INVALID, "unknown preconditioner: " + str(precond))
        d.precond = _lib.PRECOND[precond]
        if problem not in _lib.PROBLEM:
            raise PfemError(_lib.PFEM_E_INVALID, "unknown problem: " + str(problem))
        d.problem, d.tau0, d.m_reg = _lib.PROBLEM[problem], tau0, mReg
        self.ctx = ctx
        self.dim = dim
        self.kind = kind
        self.desc = d
        self.wc_state = wc_state
        self._h = C.c_void_p()
        check(lib().pfem_case_create(ctx._h, C.byref(d), C.byref(self._h)))
        self.N = 0
        self.E = 0

    def close(self):
        if self._h:
            lib().pfem_case_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- mesh and state --------------------------------------------------------------
    def set_mesh(self, mesh: PfemMesh):
        if mesh.dim != self.dim:
            raise PfemError(_lib.PFEM_E_INVALID, "mesh dimension differs from the case's")
        conn = np.ascontiguousarray(mesh.conn, dtype=np.int32)
        flags = np.ascontiguousarray(mesh.flags, dtype=np.uint8)
        pos = np.ascontiguousarray(mesh.pos, dtype=np.float64)
        check(lib().pfem_case_set_mesh(self._h, mesh.n_nodes, mesh.n_elems, _ip(conn), _u8(flags), _dp(pos)))
        self.N = mesh.n_nodes
        self.E = mesh.n_elems

    def set_mesh_from(self, mesher: "Mesher"):
        """The mesher's current mesh and node states (after Mesher.triangulate)."""
        check(lib().pfem_case_set_mesh_from(self._h, mesher._h))
        self.N, self.E, _ = mesher.sizes()

    def engine_stats(self) -> dict:
        """Diagnostics of the case's own engine: ticks, graph replays and captures, peak in flight."""
        o = (C.c_int64 * 4)()
        check(lib().pfem_case_engine_stats(self._h, o))
        return dict(ticks=o[0], graph_launches=o[1], graph_captures=o[2], peak_inflight=o[3])

    def setup_times(self) -> dict:
        """The last set_mesh: seconds in the pattern, ELL structure, multigrid hierarchy and in
        total, the smallest element shape quality (3D) and the smoother damping it implies."""
        o = np.zeros(6)
        check(lib().pfem_case_setup_times(self._h, _dp(o)))
        return dict(zip(["pattern", "ell", "multigrid", "total", "min_quality", "amg_omega"], o.tolist()))

    def remesh(self, mesher: "Mesher"):
        """Mesh::remesh between steps (pfem_case_remesh): positions and node states go to the
        mesher, the new mesh and states come back; Dirichlet values must be set again."""
        check(lib().pfem_case_remesh(self._h, mesher._h))
        self.N, self.E, _ = mesher.sizes()

    @property
    def n_state(self):
        if self.kind == "HEAT":
            return self.N
        if self.kind == "Boussinesq":
            return (self.dim + 2) * self.N
        return (self.dim + 1) * self.N if self.kind in ("PSPG", "FracStep") else (2 * self.dim + 2) * self.N

    def set_state(self, q):
        q = np.ascontiguousarray(q, dtype=np.float64).reshape(-1)
        if q.size != self.n_state:
            raise PfemError(_lib.PFEM_E_INVALID, f"state has {q.size} entries, expected {self.n_state}")
        check(lib().pfem_case_set_state(self._h, _dp(q)))

    def get_state(self) -> np.ndarray:
        q = np.empty(self.n_state)
        check(lib().pfem_case_get_state(self._h, _dp(q)))
        return q

    def set_positions(self, pos):
        pos = np.ascontiguousarray(pos, dtype=np.float64).reshape(-1)
        check(lib().pfem_case_set_positions(self._h, _dp(pos)))

    def get_positions(self) -> np.ndarray:
        pos = np.empty(self.dim * self.N)
        check(lib().pfem_case_get_positions(self._h, _dp(pos)))
        return pos.reshape(self.dim, self.N)

    def set_dirichlet(self, vals):
        vals = np.ascontiguousarray(vals, dtype=np.float64).reshape(-1)
        check(lib().pfem_case_set_dirichlet(self._h, _dp(vals)))

    def set_surface_facets(self, nodes, elem):
        """The mesh's boundary facets (F x dim nodes, F elements) for the surface-tension term
        (pfem_case_set_facets); set by set_mesh_from / remesh when the case has gamma > 0."""
        nodes = np.ascontiguousarray(nodes, dtype=np.int32)
        elem = np.ascontiguousarray(elem, dtype=np.int32)
        F = len(elem)
        check(lib().pfem_case_set_facets(self._h, F, _ip(nodes) if F else None, _ip(elem) if F else None))

    def footprint(self) -> int:
        b = C.c_int64(0)
        check(lib().pfem_case_footprint(self._h, C.byref(b)))
        return b.value

    def workspace_bytes(self) -> int:
        b = C.c_int64(0)
        check(lib().pfem_case_workspace_bytes(self._h, C.byref(b)))
        return b.value

    def reserve_workspace(self):
        check(lib().pfem_case_reserve_workspace(self._h))

    def release_workspace(self):
        check(lib().pfem_case_release_workspace(self._h))

    # ---- PSPG building blocks -------------------------------------------------------
    def assemble(self, dt, q_prev=None, q_iter=None):
        qp = None if q_prev is None else np.ascontiguousarray(q_prev, dtype=np.float64)
        qi = None if q_iter is None else np.ascontiguousarray(q_iter, dtype=np.float64)
        check(lib().pfem_pspg_assemble(self._h, dt, _dp(qp) if qp is not None else None,
                                       _dp(qi) if qi is not None else None))

    def solve(self, x0=None):
        x0 = None if x0 is None else np.ascontiguousarray(x0, dtype=np.float64)
        n = (self.dim + 1) * self.N
        x = np.empty(n)
        st = _lib.SolveStats()
        rc = check(lib().pfem_pspg_solve(self._h, _dp(x0) if x0 is not None else None, _dp(x), C.byref(st)),
                   allow_positive=True)
        return rc, x, st.as_dict()

    def solve_device(self):
        """pfem_pspg_solve with the device iterate as initial guess and result (nothing copied
        to the host); returns (status, None, stats)."""
        st = _lib.SolveStats()
        rc = check(lib().pfem_pspg_solve(self._h, None, None, C.byref(st)), allow_positive=True)
        return rc, None, st.as_dict()

    def csr(self):
        n = C.c_int64()
        nnz = C.c_int64()
        check(lib().pfem_get_csr_sizes(self._h, C.byref(n), C.byref(nnz)))
        rp = np.empty(n.value + 1, np.int32)
        ci = np.empty(nnz.value, np.int32)
        va = np.empty(nnz.value)
        check(lib().pfem_get_csr(self._h, _ip(rp), _ip(ci), _dp(va)))
        return rp, ci, va

    def rhs(self):
        b = np.empty((self.dim + 1) * self.N)
        check(lib().pfem_get_rhs(self._h, _dp(b)))
        return b

    def spmv(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.empty_like(x)
        check(lib().pfem_spmv(self._h, _dp(x), _dp(y)))
        return y

    def spmv_ell(self, x):
        """A x through the sliced ELL copy the BiCGSTAB kernels read (instrumentation)."""
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.empty_like(x)
        check(lib().pfem_spmv_ell(self._h, _dp(x), _dp(y)))
        return y

    # ---- the Picard loop's callbacks one by one (PicardAlgo.cpp:31-95) -----------------
    def save_positions(self):
        """Mesh::saveNodesList at the start of a step (node states = q_prev)."""
        check(lib().pfem_save_positions(self._h))

    def restore_positions(self):
        """Mesh::restoreNodesList: saved positions, node states back to q_prev."""
        check(lib().pfem_restore_positions(self._h))

    def update_positions(self, dt, q=None):
        """setNodesStatesfromQ + updateNodesPositionFromSave (q None: the device iterate)."""
        qa = None if q is None else np.ascontiguousarray(q, dtype=np.float64)
        check(lib().pfem_update_positions(self._h, dt, _dp(qa) if qa is not None else None))

    def residual(self, kind=None, q_iter=None, q_iter_prev=None) -> float:
        """Residual callback (kind "U", "U_P" or "Ax_f"; None: the case's); afterwards the
        device's previous iterate is the current one."""
        k = _lib.RES[kind] if kind is not None else self.desc.residual
        qi = None if q_iter is None else np.ascontiguousarray(q_iter, dtype=np.float64)
        qp = None if q_iter_prev is None else np.ascontiguousarray(q_iter_prev, dtype=np.float64)
        out = C.c_double(0)
        check(lib().pfem_residual(self._h, k, _dp(qi) if qi is not None else None,
                                  _dp(qp) if qp is not None else None, C.byref(out)))
        return out.value

    def commit(self):
        """q_prev = node states (the next step's qPrev)."""
        check(lib().pfem_pspg_commit(self._h))

    def iterate(self) -> np.ndarray:
        q = np.empty((self.dim + 1) * self.N)
        check(lib().pfem_get_iterate(self._h, _dp(q)))
        return q

    # ---- time steps ---------------------------------------------------------------------
    def pspg_step(self, dt):
        st = _lib.StepStats()
        rc = check(lib().pfem_pspg_step(self._h, dt, C.byref(st)), allow_positive=True)
        return rc, st.as_dict()

    def wc_step(self, dt):
        st = _lib.StepStats()
        check(lib().pfem_wc_step(self._h, dt, C.byref(st)))
        return st.as_dict()

    def profile_spmv(self, enable: bool):
        """Start (True) or stop (False) live event timing of the BiCGSTAB SpMV kernel; returns
        {avg_ms_fp64, launches_fp64, avg_ms_fp32, launches_fp32} gathered since the last call."""
        out = np.zeros(4)
        check(lib().pfem_profile_spmv(self._h, 1 if enable else 0, _dp(out)))
        return dict(avg_ms_fp64=out[0], launches_fp64=int(out[1]), avg_ms_fp32=out[2], launches_fp32=int(out[3]))

    def bench_kernels(self, reps=20):
        """Back-to-back timings (microseconds per launch) on the assembled system: the FP64
        BiCGSTAB SpMV kernel as a solve launches it, the FP32 node-block SpMV, a streaming read
        of the nnz FP64 values and a copy kernel, with the copy's bytes, nnz, node-graph nnz
        and lanes per node; then the multigrid's level-0 residual kernel (0 without a
        multigrid) and its workgroup count."""
        out = np.zeros(10)
        check(lib().pfem_bench_kernels(self._h, int(reps), _dp(out)))
        keys = ("kry_v_us", "spmv32_us", "stream_read_us", "copy_us", "copy_bytes", "nnz", "graph_nnz", "lanes",
                "amg_res0_us", "amg_res0_blocks")
        return dict(zip(keys, out.tolist()))

    def wc_next_dt(self, security_coeff, max_dt):
        out = C.c_double(0)
        check(lib().pfem_wc_next_dt(self._h, security_coeff, max_dt, C.byref(out)))
        return out.value

    def mass(self) -> float:
        """The problem's global "mass" (MassExtractor; Problem::getGlobalWrittableData)."""
        out = C.c_double(0)
        check(lib().pfem_case_mass(self._h, C.byref(out)))
        return out.value

    def minmax(self, coordinate: int, kind: str = "max") -> float:
        """MinMaxExtractor: smallest or largest coordinate of the nodes neither bound nor free."""
        if kind not in ("min", "max"):
            raise PfemError(_lib.PFEM_E_INVALID, "kind must be 'min' or 'max'")
        out = C.c_double(0)
        check(lib().pfem_case_minmax(self._h, int(coordinate), 1 if kind == "max" else 0, C.byref(out)))
        return out.value


class MomContEqIncompNewton(Case):
    """PSPG momentum-continuity equation (MomContEquation.inl) on the GPU.

    ``solve(dt)`` runs PicardAlgo::solve for one time step and returns True on success
    and False on numerical failure (solver failure, iterCount > maxIter, NaN residual),
    after which positions and states are restored, as in the reference."""

    def __init__(self, ctx, dim, **kw):
        super().__init__(ctx, dim, kind="PSPG", **kw)
        self.last_stats = None

    def solve(self, dt):  # noqa: D401  (mirrors Equation::solve)
        rc, st = self.pspg_step(dt)
        self.last_stats = st
        return rc == _lib.PFEM_OK


class MomContEqFracStep(Case):
    """The momentum-continuity equation with Solver.id = "FracStep" (MomContEquationFracStep.inl)
    on the GPU: keywords of the MomContEq table (residual "U" or "U_P", maxIter, minRes, gammaFS)
    and of the Material table (rho, mu; problem "Bingham" with tau0, mReg).  ``solve(dt)`` runs
    PicardAlgo::solve with the three-solve callbacks for one time step and returns True on
    success, False on a numerical failure (positions and states restored), like the reference."""

    def __init__(self, ctx, dim, gammaFS=1.0, residual="U", krylov_rtol=1e-12, krylov_max_iter=100000, **kw):
        super().__init__(ctx, dim, kind="FracStep", gammaFS=gammaFS, residual=residual, krylov_rtol=krylov_rtol,
                         krylov_max_iter=krylov_max_iter, **kw)
        self.last_stats = None

    def step(self, dt):
        """(status, stats dict) of one time step."""
        st = _lib.StepStats()
        rc = check(lib().pfem_frac_step(self._h, float(dt), C.byref(st)), allow_positive=True)
        self.last_stats = st.as_dict()
        return rc, self.last_stats

    def solve(self, dt) -> bool:  # noqa: D401  (mirrors Equation::solve)
        rc, _ = self.step(dt)
        return rc == _lib.PFEM_OK


class PicardAlgo:
    """PicardAlgo::solve (nonLinearAlgo/PicardAlgo.cpp:31-95) on the host, driving a PSPG case
    through the three callbacks of MomContEquationPSPG.inl:262-372 one C-ABI call at a time,
    as the reference's own PicardAlgo would bind them:

    * prepare: save the nodes, assemble and apply the BCs with qPrev;
    * solve: solve, set the node states and move the nodes from the saved positions,
      re-evaluate the Dirichlet values at the new positions (``dirichlet(pos)``, the Lua
      "<Tag>V"(pos, t + dt) call of MomContEquationPSPG.inl:206-214), re-assemble;
    * residual: the case's residual kind.

    ``dirichlet`` (optional) maps the positions (dim x N) to the Dirichlet values (dim x N);
    it is called before every assembly.  Returns True on success, False on numerical
    failure (nodes restored), like the reference."""

    def __init__(self, case: Case, max_iter: int | None = None, min_res: float | None = None, dirichlet=None):
        self.case = case
        self.max_iter = case.desc.max_iter if max_iter is None else max_iter
        self.min_res = case.desc.min_res if min_res is None else min_res
        self.dirichlet = dirichlet
        self.iterations = 0
        self.residuals = []

    def _bc(self):
        if self.dirichlet is not None:
            self.case.set_dirichlet(self.dirichlet(self.case.get_positions()))

    def solve(self, dt: float) -> bool:
        c = self.case
        self.residuals = []
        # prepare
        c.save_positions()
        self._bc()
        c.assemble(dt)
        it = 0
        res = float(np.finfo(np.float64).max)
        while res > self.min_res:
            if it > self.max_iter:
                c.restore_positions()
                self.iterations = it
                return False
            # solve callback
            rc, _, _ = c.solve_device()
            if rc != _lib.PFEM_OK:
                c.restore_positions()
                self.iterations = it
                return False
            c.update_positions(dt)
            self._bc()
            c.assemble(dt)
            # residual callback
            res = c.residual()
            self.residuals.append(res)
            if np.isnan(res):
                c.restore_positions()
                self.iterations = it + 1
                return False
            it += 1
        c.commit()
        self.iterations = it
        return True


class _HeatBC:
    """The heat equation's boundary data and solve on a case that holds it (HEAT, Boussinesq)."""

    def set_facets(self, nodes, out_node, kinds):
        nodes = np.ascontiguousarray(nodes, dtype=np.int32)
        out_node = np.ascontiguousarray(out_node, dtype=np.int32)
        kinds = np.ascontiguousarray(kinds, dtype=np.uint8)
        F = len(out_node)
        check(lib().pfem_heat_set_facets(self._h, F, _ip(nodes) if F else None, _ip(out_node) if F else None,
                                         _u8(kinds) if F else None))
        self.F = F

    def facet_points(self) -> np.ndarray:
        out = np.empty((max(self.F, 1), 3, self.dim))
        check(lib().pfem_heat_facet_points(self._h, _dp(out)))
        return out[: self.F]

    def set_flux(self, q=None):
        if q is None:
            check(lib().pfem_heat_set_flux(self._h, None))
            return
        q = np.ascontiguousarray(q, dtype=np.float64).reshape(-1)
        if q.size != self.F * 3 * self.dim:
            raise PfemError(_lib.PFEM_E_INVALID, "flux must hold F x 3 x dim values")
        check(lib().pfem_heat_set_flux(self._h, _dp(q)))

    def set_dirichlet_T(self, t_bc):
        t = np.ascontiguousarray(t_bc, dtype=np.float64).reshape(-1)
        if t.size != self.N:
            raise PfemError(_lib.PFEM_E_INVALID, "Dirichlet temperatures must hold N values")
        check(lib().pfem_heat_set_dirichlet(self._h, _dp(t)))

    def heat_assemble(self, dt):
        check(lib().pfem_heat_assemble(self._h, float(dt)))

    def system(self):
        """(row_ptr, col_idx, vals, b) of the last assembled heat system."""
        nnz = C.c_int64()
        check(lib().pfem_heat_system_nnz(self._h, C.byref(nnz)))
        rp = np.empty(self.N + 1, np.int32)
        ci = np.empty(nnz.value, np.int32)
        v = np.empty(nnz.value)
        b = np.empty(self.N)
        check(lib().pfem_heat_get_system(self._h, _ip(rp), _ip(ci), _dp(v), _dp(b)))
        return rp, ci, v, b

    def heat_step(self, dt):
        """(status, stats dict) of one HeatEqIncompNewton::solve."""
        st = _lib.StepStats()
        rc = check(lib().pfem_heat_step(self._h, float(dt), C.byref(st)), allow_positive=True)
        return rc, st.as_dict()


class HeatEqIncompNewton(_HeatBC, Case):
    """HeatEqIncompNewton without phase change (IncompNewton/HeatEquation.inl) as the Conduction
    problem runs it: one node state T, fixed nodes.  Material keywords follow the Lua Material
    table (k, rho, cv, h, Tinf, epsRad); the facets and their BC kinds come from the mesher
    (set_facets), the Dirichlet temperatures and the flux at the facet Gauss points from the
    caller (set_dirichlet_T, set_flux).  ``solve(dt)`` returns True on success, False on a
    solver failure, like Equation::solve."""

    def __init__(self, ctx, dim, rho=1000.0, cv=4180.0, k=0.6, h=0.0, Tinf=0.0, epsRad=0.0, krylov_rtol=1e-12,
                 krylov_max_iter=100000, **kw):
        super().__init__(ctx, dim, kind="HEAT", rho=rho, cv=cv, k=k, h=h, Tinf=Tinf, epsRad=epsRad,
                         krylov_rtol=krylov_rtol, krylov_max_iter=krylov_max_iter, **kw)
        self.last_stats = None
        self.F = 0

    def assemble(self, dt):  # noqa: D401
        self.heat_assemble(dt)

    def solve(self, dt) -> bool:  # noqa: D401  (mirrors Equation::solve)
        rc, self.last_stats = self.heat_step(dt)
        return rc == _lib.PFEM_OK


class BoussinesqIncompNewton(_HeatBC, Case):
    """The "Boussinesq" problem of SolverIncompNewton (IncompNewton/Solver.cpp:58-130, 249-283): the
    PSPG momentum-continuity equation whose body force F and PSPG term H are scaled by
    1 - alpha (T - Tr) at each Gauss point (MomContEquation.inl:164-197), and the heat equation
    (no phase change), on one mesh.  Node states (v, p, T), (dim + 2) N.  Velocity Dirichlet
    nodes carry PFEM_NODE_DIRICHLET (values: set_dirichlet), temperature Dirichlet nodes
    PFEM_NODE_DIRICHLET_T (values: set_dirichlet_T); the heat facets as for
    HeatEqIncompNewton.  Keywords of the Material table (rho, mu, k, cv, alpha, Tr, h, Tinf,
    epsRad; gamma and DgammaDT must be 0) and of Solver (solveHeatFirst) and MomContEq."""

    def __init__(self, ctx, dim, solve_heat_first=True, rho=1000.0, mu=1e-3, cv=4180.0, k=0.6, alpha=0.0, Tr=0.0,
                 h=0.0, Tinf=0.0, epsRad=0.0, heat_rtol=1e-12, **kw):
        super().__init__(ctx, dim, kind="Boussinesq", rho=rho, mu=mu, cv=cv, k=k, alpha=alpha, Tr=Tr, h=h, Tinf=Tinf,
                         epsRad=epsRad, **kw)
        self.solve_heat_first = solve_heat_first
        self.heat_rtol = heat_rtol
        self.F = 0
        self.last_stats = None
        self.failed = ""

    def solve(self, dt) -> bool:  # noqa: D401  (m_solveBoussinesq without the remesh)
        """Heat and momentum in the solveHeatFirst order; False on the first failure (``failed``
        names the equation; a momentum failure after the heat step keeps the new temperatures,
        as the reference's node list saved after the heat solve does)."""
        self.failed = ""
        order = ("heat", "momentum") if self.solve_heat_first else ("momentum", "heat")
        stats = {}
        for eq in order:
            rc, st = self.heat_step(dt) if eq == "heat" else self.pspg_step(dt)
            stats[eq] = st
            if rc != _lib.PFEM_OK:
                self.failed = eq
                self.last_stats = stats
                return False
        self.last_stats = stats
        return True


class SolverIncompNewton:
    """SolverIncompNewton with Solver.id = "PSPG" (srcs/simulation/physics/IncompNewton/Solver.cpp):
    solveOneTimeStep = one PSPG time step, then, on success, the time advances and the mesh is
    remeshed (m_solveIncompNewtonNoT, Solver.cpp:236-246); computeNextDT (Solver.cpp:217-232)
    divides dt by coeffDTDecrease after a failure and multiplies it by coeffDTincrease (capped
    at maxDT) after a success when adaptDT, and raises on a failure otherwise.

    The step is the fused pfem_pspg_step, or PicardAlgo when ``dirichlet(pos)`` is given (it is
    then re-evaluated before every assembly and after every remesh, as the reference's Lua
    BC function is)."""

    def __init__(self, case: Case, mesher: Mesher, initial_dt: float, max_dt: float, adapt_dt: bool = True,
                 coeff_dt_increase: float = 1.5, coeff_dt_decrease: float = 2.0, dirichlet=None, remesh: bool = True):
        self.case, self.mesher = case, mesher
        self.dt = initial_dt
        self.max_dt, self.adapt_dt = max_dt, adapt_dt
        self.coeff_inc, self.coeff_dec = coeff_dt_increase, coeff_dt_decrease
        self.dirichlet = dirichlet
        self.remesh_enabled = remesh
        self.time = 0.0
        self.succeeded = False
        self.picard = PicardAlgo(case, dirichlet=dirichlet) if dirichlet is not None else None

    def solveOneTimeStep(self) -> bool:  # noqa: N802  (reference name)
        if self.picard is not None:
            ok = self.picard.solve(self.dt)
        else:
            rc, _ = self.case.pspg_step(self.dt)
            ok = rc == _lib.PFEM_OK
        self.succeeded = ok
        if ok:
            self.time += self.dt
            if self.remesh_enabled:
                self.case.remesh(self.mesher)
                if self.dirichlet is not None:
                    self.case.set_dirichlet(self.dirichlet(self.case.get_positions()))
        return ok

    def computeNextDT(self):  # noqa: N802
        if not self.succeeded:
            if not self.adapt_dt:
                raise PfemError(_lib.PFEM_NOT_CONVERGED, "solving the precedent time step was not successful!")
            self.dt /= self.coeff_dec
        elif self.adapt_dt:
            self.dt = min(self.max_dt, self.dt * self.coeff_inc)
        return self.dt


class WCompNewtonStep(Case):
    """Weakly compressible CDS_dpdt solver step (WCompNewton/Solver.cpp)."""

    def __init__(self, ctx, dim, **kw):
        super().__init__(ctx, dim, kind="WC", **kw)

    def solveOneTimeStep(self, dt):  # noqa: N802  (reference name)
        self.wc_step(dt)
        return True

    def computeNextDT(self, security_coeff, max_dt):  # noqa: N802
        return self.wc_next_dt(security_coeff, max_dt)


def batch_step(ctx: Context, cases, dts):
    """One time step of every case as one batch (pfem_batch_step); returns (status, stats)."""
    n = len(cases)
    hs = (C.c_void_p * n)(*[c._h.value for c in cases])
    d = np.ascontiguousarray(dts, dtype=np.float64)
    status = np.zeros(n, np.int32)
    stats = (_lib.StepStats * n)()
    check(lib().pfem_batch_step(ctx._h, n, hs, _dp(d), _ip(status), stats))
    return status, [s.as_dict() for s in stats]


def now_ns() -> int:
    """The serving clock (pfem_now_ns, monotonic nanoseconds)."""
    return int(lib().pfem_now_ns())


class Server:
    """Continuous-batching request loop (pfem_server_*): one request is one time step of one
    case; results come back through ``poll``."""

    def __init__(self, ctx: Context, hbm_budget: int = 0, max_inflight: int = 64, max_chunk: int = 0,
                 wc_security_coeff: float = 0.5, wc_max_dt: float = 1.0):
        d = _lib.ServerDesc()
        d.hbm_budget, d.max_inflight, d.max_chunk = int(hbm_budget), int(max_inflight), int(max_chunk)
        d.wc_security_coeff, d.wc_max_dt = wc_security_coeff, wc_max_dt
        self._h = C.c_void_p()
        check(lib().pfem_server_create(ctx._h, C.byref(d), C.byref(self._h)))
        self._buf = (_lib.ServerResult * 256)()

    def submit(self, case: Case, dt: float, tag: int, arrival_ns: int = 0):
        check(lib().pfem_server_submit(self._h, case._h, dt, int(tag), int(arrival_ns)))

    def poll(self, max_results: int = 256, timeout_ms: int = 0):
        """Completed results (dicts with tag, status, stats, next_dt and the timestamps)."""
        n = C.c_int32(0)
        m = min(max_results, len(self._buf))
        check(lib().pfem_server_poll(self._h, self._buf, m, int(timeout_ms), C.byref(n)))
        return [self._buf[i].as_dict() for i in range(n.value)]

    def stats(self):
        out = (C.c_int64 * 8)()
        check(lib().pfem_server_stats(self._h, out))
        keys = ["ticks", "peak_inflight", "arena_bytes", "arena_peak", "submitted", "completed", "arena_used",
                "not_arrived"]
        return dict(zip(keys, [int(v) for v in out]))

    def close(self):
        if self._h:
            lib().pfem_server_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Mesher:
    """The reference's Mesh after loadFromFile (srcs/mesh/Mesh.hpp): alpha-shape triangulation
    and remeshing of a node cloud on the host (pfem_mesher_*).  Keyword names follow the Lua
    Mesh table (Problem.cpp:56-70): hchar, alpha, omega, gamma, boundingBox, exclusionZones,
    addOnFS, deleteFlyingNodes, laplacianSmoothingBoundaries."""

    def __init__(self, dim: int, hchar: float, alpha: float = 1.0, omega: float = 1e16, gamma: float = 0.0,
                 boundingBox=None, exclusionZones=(), addOnFS: bool = True, deleteFlyingNodes: bool = False,
                 laplacianSmoothingBoundaries: bool = False):
        d = _lib.MesherDesc()
        d.dim, d.hchar, d.alpha, d.omega, d.gamma = dim, hchar, alpha, omega, gamma
        d.add_on_fs, d.delete_flying_nodes = int(addOnFS), int(deleteFlyingNodes)
        d.laplacian_smoothing_boundaries = int(laplacianSmoothingBoundaries)
        if boundingBox is not None:
            if len(boundingBox) != 2 * dim:
                raise PfemError(_lib.PFEM_E_INVALID, f"Invalid bounding box size: {len(boundingBox)}")
            d.has_bounding_box = 1
            for k, v in enumerate(boundingBox):
                d.bounding_box[k] = float(v)
        zones = np.ascontiguousarray(np.asarray(exclusionZones, dtype=np.float64).reshape(-1, 2 * dim))
        self._zones = zones
        d.n_exclusion_zones = zones.shape[0]
        d.exclusion_zones = _dp(zones) if zones.shape[0] else None
        self.dim = dim
        self._h = C.c_void_p()
        check(lib().pfem_mesher_create(C.byref(d), C.byref(self._h)))
        self.n_states = 0

    def close(self):
        if self._h:
            lib().pfem_mesher_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_nodes(self, pos, flags, states=None, tags=None):
        pos = np.ascontiguousarray(pos, dtype=np.float64)
        n = pos.shape[1]
        flags = np.ascontiguousarray(flags, dtype=np.uint8)
        ns = 0
        st = None
        if states is not None:
            st = np.ascontiguousarray(states, dtype=np.float64).reshape(-1)
            ns = st.size // n
        tg = None if tags is None else np.ascontiguousarray(tags, dtype=np.int32)
        check(lib().pfem_mesher_set_nodes(self._h, n, ns, _dp(pos), _dp(st) if st is not None else None, _u8(flags),
                                          _ip(tg) if tg is not None else None))
        self.n_states = ns

    def set_positions_states(self, pos=None, states=None):
        p = None if pos is None else np.ascontiguousarray(pos, dtype=np.float64)
        s = None if states is None else np.ascontiguousarray(states, dtype=np.float64)
        check(lib().pfem_mesher_set_positions_states(self._h, _dp(p) if p is not None else None,
                                                     _dp(s) if s is not None else None))

    def triangulate(self):
        check(lib().pfem_mesher_triangulate(self._h))

    def remesh(self):
        check(lib().pfem_mesher_remesh(self._h))

    def sizes(self):
        n, e, f = C.c_int32(), C.c_int32(), C.c_int32()
        check(lib().pfem_mesher_sizes(self._h, C.byref(n), C.byref(e), C.byref(f)))
        return n.value, e.value, f.value

    def nodes(self):
        """(pos dim x N, states nstates x N, flags N, tags N)."""
        n, _, _ = self.sizes()
        pos = np.empty((self.dim, n))
        st = np.empty(max(self.n_states, 1) * n)
        fl = np.empty(n, dtype=np.uint8)
        tg = np.empty(n, dtype=np.int32)
        check(lib().pfem_mesher_get_nodes(self._h, _dp(pos), _dp(st), _u8(fl), _ip(tg)))
        return pos, st[: self.n_states * n].reshape(self.n_states, n), fl, tg

    def elements(self) -> np.ndarray:
        _, e, _ = self.sizes()
        conn = np.empty((e, self.dim + 1), dtype=np.int32)
        check(lib().pfem_mesher_get_elements(self._h, _ip(conn)))
        return conn

    def facets(self):
        """(nodes F x dim, out node F, element F) of the free-surface facets."""
        _, _, f = self.sizes()
        nodes = np.empty((f, self.dim), dtype=np.int32)
        out = np.empty(f, dtype=np.int32)
        el = np.empty(f, dtype=np.int32)
        check(lib().pfem_mesher_get_facets(self._h, _ip(nodes), _ip(out), _ip(el)))
        return nodes, out, el

    def stats(self) -> dict:
        o = np.zeros(8)
        check(lib().pfem_mesher_stats(self._h, _dp(o)))
        keys = ["added", "removed", "out_of_box", "flying_deleted", "duplicates", "t_nodes", "t_delaunay", "t_alpha"]
        return dict(zip(keys, o.tolist()))

    def mesh(self, h: float | None = None, name: str = "") -> PfemMesh:
        """The current triangulation as a PfemMesh (flags include ONFS and FREE)."""
        pos, _, fl, _ = self.nodes()
        return PfemMesh(self.dim, pos, self.elements(), fl, h if h is not None else 0.0, name=name)
