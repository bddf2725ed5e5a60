"""ctypes front-end of the CPU oracle (liboracle.so built from oracle/pfem_oracle.cpp).

TEST INFRASTRUCTURE: imported only by tests/, __graft_entry__.smoke() and the CPU-baseline
leg of bench.py, always as the parity reference or the timed CPU baseline — never by the product
path (pfem_amd/).  The restatement it wraps follows the reference files cited in
pfem_oracle.cpp; the pins that check it live in tests/test_oracle_*.py.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# PFEM_ORACLE_LIB points the wrapper at another build of the same sources (the mutation test
# builds deliberately broken copies to show that the pins catch them)
_LIB_PATH = os.environ.get("PFEM_ORACLE_LIB") or os.path.join(_HERE, "liboracle.so")
_lib = None

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)
_u8p = C.POINTER(C.c_uint8)


def build(force: bool = False) -> str:
    """Compile liboracle.so in place (make -C oracle)."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-C", _HERE, "-B" if force else "-s"], check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        L.orc_pspg_assemble.restype = C.c_void_p
        L.orc_pspg_assemble.argtypes = [C.c_int, C.c_int64, C.c_int64, _ip, _u8p, _dp, _dp, _dp, _dp, C.c_double,
                                        C.c_double, _dp, C.c_double, C.c_int, C.c_int]
        L.orc_sys_n.restype = C.c_int64
        L.orc_sys_n.argtypes = [C.c_void_p]
        L.orc_sys_nnz.restype = C.c_int64
        L.orc_sys_nnz.argtypes = [C.c_void_p]
        L.orc_sys_copy.argtypes = [C.c_void_p, _ip, _ip, _dp, _dp]
        L.orc_sys_times.argtypes = [C.c_void_p, _dp]
        L.orc_sys_free.argtypes = [C.c_void_p]
        L.orc_pspg_element.argtypes = [C.c_int, _dp, _dp, _dp, C.c_double, C.c_double, _dp, C.c_double, _dp, _dp, _dp]
        L.orc_element_matrices.argtypes = [C.c_int, _dp, C.c_double, C.c_double, _dp, _dp, _dp, _dp, _dp, _dp, _dp,
                                           _dp]
        L.orc_sys_solve.restype = C.c_int
        L.orc_sys_solve.argtypes = [C.c_void_p, C.c_int, C.c_double, C.c_int, _dp, _ip, _dp, C.c_int]
        L.orc_csr_bicgstab.restype = C.c_int
        L.orc_csr_bicgstab.argtypes = [C.c_int64, _ip, _ip, _dp, _dp, _dp, C.c_int, C.c_double, C.c_int, _ip, _dp,
                                       C.c_int]
        L.orc_pspg_step.restype = C.c_int
        L.orc_pspg_step.argtypes = [C.c_int, C.c_int64, C.c_int64, _ip, _u8p, _dp, _dp, _dp, C.c_double, C.c_double,
                                    _dp, C.c_double, C.c_int, C.c_double, C.c_int, C.c_int, C.c_double, C.c_int, _dp,
                                    C.c_int, _dp]
        L.orc_wc_step.restype = C.c_int
        L.orc_wc_step.argtypes = [C.c_int, C.c_int64, C.c_int64, _ip, _u8p, _dp, _dp, _dp, _dp, _dp, _dp, C.c_double,
                                  _dp, C.c_double, C.c_double, C.c_double, C.c_int, C.c_double, _dp, C.c_int]
        L.orc_wc_next_dt.restype = C.c_double
        L.orc_wc_next_dt.argtypes = [C.c_int, C.c_int64, C.c_int64, _ip, _dp, _dp, _dp, _dp, C.c_double, C.c_double,
                                     C.c_double, C.c_double, C.c_double, C.c_int]
        L.orc_mass.restype = C.c_double
        L.orc_mass.argtypes = [C.c_int, C.c_int64, C.c_int64, _ip, _dp, _dp]
        L.orc_minmax.restype = C.c_double
        L.orc_minmax.argtypes = [C.c_int, C.c_int64, _u8p, _dp, C.c_int, C.c_int]
        L.orc_picard_residual.restype = C.c_double
        L.orc_picard_residual.argtypes = [C.c_int, C.c_int64, _u8p, C.c_int, _ip, _ip, _dp, _dp, _dp, _dp]
        L.orc_set_material.argtypes = [C.c_int, C.c_double, C.c_double]
        L.orc_set_boussinesq.argtypes = [C.c_double, C.c_double, _dp, C.c_int64]
        L.orc_set_facets.argtypes = [C.c_int64, C.c_int, _ip, _ip, C.c_double, C.c_double]
        L.orc_wc_parts.argtypes = [C.c_int, C.c_int64, C.c_int64, _ip, _dp, _dp, _dp, _dp, C.c_double, _dp,
                                   C.c_double, C.c_double, C.c_int, C.c_double, _dp, _dp, _dp, _dp, C.c_int]
        L.orc_heat_assemble.restype = C.c_void_p
        L.orc_heat_assemble.argtypes = [C.c_int, C.c_int64, C.c_int64, _ip, _u8p, _dp, _dp, _dp, C.c_double,
                                        C.c_double, C.c_double, C.c_double, C.c_double, C.c_double, C.c_double,
                                        C.c_int64, _ip, _ip, _u8p, _dp]
        L.orc_heat_step.restype = C.c_int
        L.orc_heat_step.argtypes = [C.c_int, C.c_int64, C.c_int64, _ip, _u8p, _dp, _dp, _dp, C.c_double, C.c_double,
                                    C.c_double, C.c_double, C.c_double, C.c_double, C.c_double, C.c_int64, _ip, _ip,
                                    _u8p, _dp, C.c_double, C.c_int, _ip, _dp]
        L.orc_heat_facet_terms.argtypes = [C.c_int, C.c_int64, _dp, _dp, C.c_int64, _ip, _ip, _dp, C.c_double,
                                           C.c_double, C.c_double, _dp, _dp, _dp]
        L.orc_frac_step.restype = C.c_int
        L.orc_frac_step.argtypes = [C.c_int, C.c_int64, C.c_int64, _ip, _u8p, _dp, _dp, _dp, C.c_double, C.c_double,
                                    _dp, C.c_double, C.c_double, C.c_int, C.c_double, C.c_int, C.c_double, C.c_int,
                                    _dp]
        _lib = L
    return _lib


def _d(a):
    return a.ctypes.data_as(_dp)


def _i(a):
    return a.ctypes.data_as(_ip)


def _u8(a):
    return a.ctypes.data_as(_u8p)


def _g3(g):
    out = np.zeros(3)
    out[:len(g)] = g
    return out


@dataclass
class CsrSystem:
    n: int
    row_ptr: np.ndarray
    col_idx: np.ndarray
    vals: np.ndarray
    b: np.ndarray
    times: np.ndarray  # Compute triplets, Push back, Assemble matrix, Assemble vector, Apply BC

    def to_scipy(self):
        import scipy.sparse as sp
        return sp.csr_matrix((self.vals, self.col_idx, self.row_ptr), shape=(self.n, self.n))


def pspg_assemble(mesh, q_prev, q_iter=None, dir_vals=None, rho=1000.0, mu=1e-3, g=None, dt=1e-3, apply_bc=True,
                  nthreads=0) -> CsrSystem:
    """m_buildAbPSPG (+ m_applyBCPSPG): MomContEquationPSPG.inl:7-235."""
    L = lib()
    d = mesh.dim
    N = mesh.n_nodes
    g = _g3(g if g is not None else ([0.0, -9.81] if d == 2 else [0.0, 0.0, -9.81]))
    q_prev = np.ascontiguousarray(q_prev, dtype=np.float64)
    q_iter = q_prev if q_iter is None else np.ascontiguousarray(q_iter, dtype=np.float64)
    dir_vals = np.zeros(d * N) if dir_vals is None else np.ascontiguousarray(dir_vals, dtype=np.float64)
    conn = np.ascontiguousarray(mesh.conn, dtype=np.int32)
    flags = np.ascontiguousarray(mesh.flags, dtype=np.uint8)
    pos = np.ascontiguousarray(mesh.pos, dtype=np.float64)
    h = L.orc_pspg_assemble(d, N, mesh.n_elems, _i(conn), _u8(flags), _d(pos), _d(q_prev), _d(q_iter), _d(dir_vals),
                            rho, mu, _d(g), dt, 1 if apply_bc else 0, nthreads)
    try:
        n = L.orc_sys_n(h)
        nnz = L.orc_sys_nnz(h)
        rp = np.empty(n + 1, np.int32)
        ci = np.empty(nnz, np.int32)
        v = np.empty(nnz)
        b = np.empty(n)
        L.orc_sys_copy(h, _i(rp), _i(ci), _d(v), _d(b))
        t = np.empty(5)
        L.orc_sys_times(h, _d(t))
    finally:
        L.orc_sys_free(h)
    return CsrSystem(n, rp, ci, v, b, t)


def pspg_element(dim, pos_e, q_prev_e, q_iter_e, rho=1000.0, mu=1e-3, g=None, dt=1e-3):
    """Ae, be, tau of a standalone element (PSPG.inl:26-53) with nodes 0..dim."""
    L = lib()
    na = (dim + 1) * (dim + 1)
    g = _g3(g if g is not None else ([0.0, -9.81] if dim == 2 else [0.0, 0.0, -9.81]))
    Ae = np.empty((na, na))
    be = np.empty(na)
    tau = np.empty(1)
    L.orc_pspg_element(dim, _d(np.ascontiguousarray(pos_e, dtype=np.float64)),
                       _d(np.ascontiguousarray(q_prev_e, dtype=np.float64)),
                       _d(np.ascontiguousarray(q_iter_e, dtype=np.float64)), rho, mu, _d(g), dt, _d(Ae), _d(be),
                       _d(tau))
    return Ae, be, float(tau[0])


def element_matrices(dim, pos_e, rho=1.0, mu=1.0):
    """M (factor rho), K (mu, ddev = diag(2..,1..)), D, C, L (factor 1), gradN, detJ, r_in."""
    L = lib()
    ne = dim + 1
    nd = dim * ne
    M = np.empty((ne, ne))
    K = np.empty((nd, nd))
    D = np.empty((ne, nd))
    Cm = np.empty((ne, nd))
    Lm = np.empty((ne, ne))
    G = np.empty((dim, ne))
    dj = np.empty(1)
    rin = np.empty(1)
    L.orc_element_matrices(dim, _d(np.ascontiguousarray(pos_e, dtype=np.float64)), rho, mu, _d(M), _d(K), _d(D),
                           _d(Cm), _d(Lm), _d(G), _d(dj), _d(rin))
    return dict(M=M, K=K, D=D, C=Cm, L=Lm, gradN=G, detJ=float(dj[0]), rin=float(rin[0]))


def solve(sys: CsrSystem, kind=2, rtol=1e-10, maxit=20000, x0=None, nthreads=0):
    """kind 0 dense LU (SciPy), 1 BiCGSTAB+Jacobi, 2 BiCGSTAB+ILU(0). Returns (status, x, iters, relres)."""
    x = np.zeros(sys.n) if x0 is None else np.array(x0, dtype=np.float64)
    if kind == 0:
        import scipy.linalg as sla
        A = sys.to_scipy().toarray()
        x[:] = sla.solve(A, sys.b)
        rr = float(np.linalg.norm(sys.b - A @ x) / max(np.linalg.norm(sys.b), 1e-300))
        return 0, x, 0, rr
    st, it, rr = csr_bicgstab(sys.n, sys.row_ptr, sys.col_idx, sys.vals, sys.b, x, 0 if kind == 1 else 1, rtol, maxit,
                              nthreads)
    return st, x, it, rr


def csr_bicgstab(n, rp, ci, v, b, x, prec=0, rtol=1e-10, maxit=1000, nthreads=0):
    L = lib()
    it = C.c_int(0)
    rr = C.c_double(0)
    st = L.orc_csr_bicgstab(n, _i(rp), _i(ci), _d(v), _d(b), _d(x), prec, rtol, maxit, C.byref(it), C.byref(rr),
                            nthreads)
    return st, it.value, rr.value


RES_KIND = {"U": 0, "U_P": 1, "Ax_f": 2}


def pspg_step(mesh, pos, q, dir_vals=None, rho=1000.0, mu=1e-3, g=None, dt=1e-3, residual="Ax_f", min_res=1e-6,
              max_iter=10, solver_kind=2, rtol=1e-10, krylov_maxit=20000, nthreads=0, dir_affine=None):
    """One Picard time step (PicardAlgo.cpp:31-95 + PSPG.inl:262-372); pos, q updated in place.
    dir_affine (dim x (dim+1)): Dirichlet values v_bc(x) = A x + c re-evaluated at the current
    positions before every assembly (instead of the constant dir_vals)."""
    L = lib()
    d = mesh.dim
    N = mesh.n_nodes
    g = _g3(g if g is not None else ([0.0, -9.81] if d == 2 else [0.0, 0.0, -9.81]))
    dir_vals = np.zeros(d * N) if dir_vals is None else np.ascontiguousarray(dir_vals, dtype=np.float64)
    stats = np.zeros(14)
    conn = np.ascontiguousarray(mesh.conn, dtype=np.int32)
    flags = np.ascontiguousarray(mesh.flags, dtype=np.uint8)
    assert pos.dtype == np.float64 and pos.flags.c_contiguous and q.dtype == np.float64 and q.flags.c_contiguous
    aff = None if dir_affine is None else np.ascontiguousarray(dir_affine, dtype=np.float64).reshape(d, d + 1)
    st = L.orc_pspg_step(d, N, mesh.n_elems, _i(conn), _u8(flags), _d(pos), _d(q), _d(dir_vals), rho, mu, _d(g), dt,
                         RES_KIND[residual], min_res, max_iter, solver_kind, rtol, krylov_maxit, _d(stats), nthreads,
                         _d(aff) if aff is not None else C.cast(None, _dp))
    keys = ["picard_iters", "krylov_iters", "assemblies", "final_res", "max_relres", "t_assemble", "t_solve",
            "t_update", "t_residual", "t_triplets", "t_push", "t_set_from_triplets", "t_vector", "t_bc"]
    return st, dict(zip(keys, stats.tolist()))


def wc_step(mesh, pos, v, p, rho, a, dt, dir_vals=None, mu=1e-3, g=None, K0=2.2e9, K0p=7.0, rho_star=1000.0,
            meduri=True, nthreads=0):
    """One CDS_dpdt step (WCompNewton/Solver.cpp:236-276); arrays updated in place."""
    L = lib()
    d = mesh.dim
    N = mesh.n_nodes
    g = _g3(g if g is not None else ([0.0, -9.81] if d == 2 else [0.0, 0.0, -9.81]))
    dir_vals = np.zeros(d * N) if dir_vals is None else np.ascontiguousarray(dir_vals, dtype=np.float64)
    for arr in (pos, v, p, rho, a):
        assert arr.dtype == np.float64 and arr.flags.c_contiguous
    t = np.zeros(2)
    conn = np.ascontiguousarray(mesh.conn, dtype=np.int32)
    flags = np.ascontiguousarray(mesh.flags, dtype=np.uint8)
    L.orc_wc_step(d, N, mesh.n_elems, _i(conn), _u8(flags), _d(pos), _d(v), _d(p), _d(rho), _d(a), _d(dir_vals), mu,
                  _d(g), K0, K0p, rho_star, 1 if meduri else 0, dt, _d(t), nthreads)
    return t


def wc_next_dt(mesh, pos, v, p, rho, mu=1e-3, K0=2.2e9, K0p=7.0, security_coeff=0.5, max_dt=1.0, nthreads=0):
    """computeNextDT (WCompNewton/Solver.cpp:192-234)."""
    L = lib()
    conn = np.ascontiguousarray(mesh.conn, dtype=np.int32)
    return L.orc_wc_next_dt(mesh.dim, mesh.n_nodes, mesh.n_elems, _i(conn), _d(pos), _d(v), _d(p), _d(rho), mu, K0,
                            K0p, security_coeff, max_dt, nthreads)


def mass(mesh, pos, rho=None):
    """Global "mass" (IncompNewton/Problem.cpp:173-190: sum of element sizes; with rho,
    WCompNewton/Problem.cpp:216-240: sizes weighted by the element's mean node density)."""
    L = lib()
    conn = np.ascontiguousarray(mesh.conn, dtype=np.int32)
    pos = np.ascontiguousarray(pos, dtype=np.float64).reshape(-1)
    r = None if rho is None else np.ascontiguousarray(rho, dtype=np.float64)
    return L.orc_mass(mesh.dim, mesh.n_nodes, mesh.n_elems, _i(conn), _d(pos),
                      _d(r) if r is not None else C.cast(None, _dp))


def minmax(mesh, pos, coordinate, kind="max", flags=None):
    """MinMaxExtractor (MinMaxExtractor.cpp:35-55) over nodes neither bound nor free; the
    library derives FREE from connectivity, so pass those flags when comparing with it."""
    L = lib()
    fl = np.ascontiguousarray(mesh.flags if flags is None else flags, dtype=np.uint8)
    pos = np.ascontiguousarray(pos, dtype=np.float64).reshape(-1)
    return L.orc_minmax(mesh.dim, mesh.n_nodes, _u8(fl), _d(pos), coordinate, 1 if kind == "max" else 0)


def picard_residual(dim, flags, kind, q_iter, q_iter_prev, system=None):
    """Residual callback of the Picard loop (MomContEquationPSPG.inl:314-372) on given iterates;
    flags are used as given (FREE included).  kind: "U", "U_P" or "Ax_f" (needs system, a
    CsrSystem)."""
    L = lib()
    k = RES_KIND[kind]
    fl = np.ascontiguousarray(flags, dtype=np.uint8)
    N = fl.size
    qi = np.ascontiguousarray(q_iter, dtype=np.float64)
    qp = np.ascontiguousarray(q_iter_prev, dtype=np.float64)
    assert qi.size == (dim + 1) * N and qp.size == qi.size
    if k == 2:
        rp, ci, va, b = system.row_ptr, system.col_idx, system.vals, system.b
        return L.orc_picard_residual(dim, N, _u8(fl), k, _i(rp), _i(ci), _d(va), _d(b), _d(qi), _d(qp))
    nul_i, nul_d = C.cast(None, _ip), C.cast(None, _dp)
    return L.orc_picard_residual(dim, N, _u8(fl), k, nul_i, nul_i, nul_d, nul_d, _d(qi), _d(qp))


def wc_parts(mesh, pos, v, p, rho, dt=1e-4, mu=1e-3, g=None, K0=2.2e9, K0p=7.0, meduri=True, nthreads=0):
    """The scattered WC systems of one step at the given state, before boundary conditions:
    continuity lumped mass ML and right-hand side F0 (ContEquation.inl:353-413), momentum
    lumped mass Mv and force F (MomEquation.inl:229-302).  Returns (ML, F0, Mv, F)."""
    L = lib()
    d, N = mesh.dim, mesh.n_nodes
    g = _g3(g if g is not None else ([0.0, -9.81] if d == 2 else [0.0, 0.0, -9.81]))
    conn = np.ascontiguousarray(mesh.conn, dtype=np.int32)
    pos = np.ascontiguousarray(pos, dtype=np.float64).reshape(-1)
    v = np.ascontiguousarray(v, dtype=np.float64).reshape(-1)
    p = np.ascontiguousarray(p, dtype=np.float64)
    rho = np.ascontiguousarray(rho, dtype=np.float64)
    ML, F0, Mv, F = np.empty(N), np.empty(N), np.empty(d * N), np.empty(d * N)
    L.orc_wc_parts(d, N, mesh.n_elems, _i(conn), _d(pos), _d(v), _d(p), _d(rho), mu, _d(g), K0, K0p,
                   1 if meduri else 0, dt, _d(ML), _d(F0), _d(Mv), _d(F), nthreads)
    return ML, F0, Mv, F


class material:
    """Context manager selecting the momentum equation's material law for the PSPG calls inside
    it (MomContEquation.inl:33-66, 100-123, 164-197): "Newtonian", "Bingham" (tau0, mReg) or
    "Boussinesq" (alpha, Tr and the node temperatures T of the F and H factors)."""

    def __init__(self, problem="Newtonian", tau0=0.0, mReg=0.0, alpha=0.0, Tr=0.0, T=None):
        self.args = ({"Newtonian": 0, "Bingham": 1, "Boussinesq": 2}[problem], float(tau0), float(mReg))
        self.bq = (float(alpha), float(Tr), None if T is None else np.ascontiguousarray(T, dtype=np.float64))

    def __enter__(self):
        lib().orc_set_material(*self.args)
        a, tr, T = self.bq
        lib().orc_set_boussinesq(a, tr, _d(T) if T is not None else C.cast(None, _dp), 0 if T is None else T.size)
        return self

    def __exit__(self, *exc):
        lib().orc_set_material(0, 0.0, 0.0)
        lib().orc_set_boussinesq(0.0, 0.0, C.cast(None, _dp), 0)
        return False


class surface_tension:
    """Context manager setting the boundary facets (nodes F x dim, element F) and gamma, DgammaDT
    of the surface-tension term (getFST, MatricesBuilder.inl:389-402) for the PSPG and WC calls
    inside it."""

    def __init__(self, nodes, elem, gamma, DgammaDT=0.0):
        self.nodes = np.ascontiguousarray(nodes, dtype=np.int32)
        self.elem = np.ascontiguousarray(elem, dtype=np.int32)
        self.gamma, self.dgdT = float(gamma), float(DgammaDT)

    def __enter__(self):
        F, d = self.nodes.shape
        lib().orc_set_facets(F, d, _i(self.nodes), _i(self.elem), self.gamma, self.dgdT)
        return self

    def __exit__(self, *exc):
        z = np.zeros(1, np.int32)
        lib().orc_set_facets(0, 2, _i(z), _i(z), 0.0, 0.0)
        return False


# ---- heat equation (HeatEquation.inl) ------------------------------------------------------
HEAT_Q, HEAT_QH, HEAT_QR = 1, 2, 4


@dataclass
class HeatCase:
    """Inputs of HeatEqIncompNewton without phase change: material (Material table: k, rho, cv,
    h, Tinf, epsRad) and the facets with their BC kinds (HEAT_* bits) and, for HEAT_Q facets,
    the "<Tag>Q" vectors at the 3 facet Gauss points (F x 3 x dim)."""
    rho: float = 1000.0
    cv: float = 4180.0
    k: float = 0.6
    h: float = 0.0
    Tinf: float = 0.0
    epsRad: float = 0.0
    facets: np.ndarray | None = None
    facet_out: np.ndarray | None = None
    facet_kind: np.ndarray | None = None
    qgp: np.ndarray | None = None


def _heat_args(mesh, hc):
    conn = np.ascontiguousarray(mesh.conn, dtype=np.int32)
    flags = np.ascontiguousarray(mesh.flags, dtype=np.uint8)
    pos = np.ascontiguousarray(mesh.pos, dtype=np.float64)
    d = mesh.dim
    F = 0 if hc.facets is None else len(hc.facets)
    fac = np.ascontiguousarray(hc.facets if F else np.zeros((1, d)), dtype=np.int32)
    out = np.ascontiguousarray(hc.facet_out if F else np.zeros(1), dtype=np.int32)
    kind = np.ascontiguousarray(hc.facet_kind if F else np.zeros(1), dtype=np.uint8)
    qgp = None if hc.qgp is None else np.ascontiguousarray(hc.qgp, dtype=np.float64)
    return conn, flags, pos, F, fac, out, kind, qgp


def heat_assemble(mesh, T, Tbc=None, hc: HeatCase = HeatCase(), dt=1e-2) -> CsrSystem:
    """m_buildAb + m_applyBC (HeatEquation.inl:209-364) from the node temperatures T."""
    L = lib()
    conn, flags, pos, F, fac, out, kind, qgp = _heat_args(mesh, hc)
    N = mesh.n_nodes
    T = np.ascontiguousarray(T, dtype=np.float64)
    Tbc = np.zeros(N) if Tbc is None else np.ascontiguousarray(Tbc, dtype=np.float64)
    h = L.orc_heat_assemble(mesh.dim, N, mesh.n_elems, _i(conn), _u8(flags), _d(pos), _d(T), _d(Tbc), hc.rho, hc.cv,
                            hc.k, hc.h, hc.Tinf, hc.epsRad, dt, F, _i(fac), _i(out), _u8(kind),
                            _d(qgp) if qgp is not None else C.cast(None, _dp))
    try:
        n = L.orc_sys_n(h)
        nnz = L.orc_sys_nnz(h)
        rp = np.empty(n + 1, np.int32)
        ci = np.empty(nnz, np.int32)
        v = np.empty(nnz)
        b = np.empty(n)
        L.orc_sys_copy(h, _i(rp), _i(ci), _d(v), _d(b))
    finally:
        L.orc_sys_free(h)
    return CsrSystem(n, rp, ci, v, b, np.zeros(5))


def heat_step(mesh, T, Tbc=None, hc: HeatCase = HeatCase(), dt=1e-2, rtol=1e-13, maxit=100000):
    """HeatEqIncompNewton::solve without phase change (one Picard pass): T updated in place.
    Returns (status, iterations, relres)."""
    L = lib()
    conn, flags, pos, F, fac, out, kind, qgp = _heat_args(mesh, hc)
    N = mesh.n_nodes
    assert T.dtype == np.float64 and T.flags.c_contiguous
    Tbc = np.zeros(N) if Tbc is None else np.ascontiguousarray(Tbc, dtype=np.float64)
    it = C.c_int(0)
    rr = C.c_double(0)
    st = L.orc_heat_step(mesh.dim, N, mesh.n_elems, _i(conn), _u8(flags), _d(pos), _d(T), _d(Tbc), hc.rho, hc.cv,
                         hc.k, hc.h, hc.Tinf, hc.epsRad, dt, F, _i(fac), _i(out), _u8(kind),
                         _d(qgp) if qgp is not None else C.cast(None, _dp), rtol, maxit, C.byref(it), C.byref(rr))
    return st, it.value, rr.value


def heat_facet_terms(mesh, T, hc: HeatCase):
    """getQN, getSGamma (convection) and getSGamma (radiation) of every facet (F x dim each)."""
    L = lib()
    conn, flags, pos, F, fac, out, kind, qgp = _heat_args(mesh, hc)
    d = mesh.dim
    qn, sgh, sgr = np.zeros((F, d)), np.zeros((F, d)), np.zeros((F, d))
    L.orc_heat_facet_terms(d, mesh.n_nodes, _d(pos), _d(np.ascontiguousarray(T, dtype=np.float64)), F, _i(fac),
                           _i(out), _d(qgp) if qgp is not None else C.cast(None, _dp), hc.h, hc.Tinf, hc.epsRad,
                           _d(qn), _d(sgh), _d(sgr))
    return qn, sgh, sgr


def frac_step(mesh, pos, q, dir_vals=None, rho=1000.0, mu=1e-3, g=None, dt=1e-3, gamma_fs=1.0, residual="U",
              min_res=1e-6, max_iter=10, rtol=1e-13, cg_maxit=0):
    """One fractional-step time step (MomContEquationFracStep.inl:7-630): pos and q = (v, p)
    updated in place on success.  Returns (status, stats) with status 0 converged, 1 Picard not
    converged, 2 NaN.  Each CG solve stops at rtol or cg_maxit iterations (2 n when 0) and is
    used as it stands, converged or not, as in the reference."""
    L = lib()
    d = mesh.dim
    N = mesh.n_nodes
    g = _g3(g if g is not None else ([0.0, -9.81] if d == 2 else [0.0, 0.0, -9.81]))
    dir_vals = np.zeros(d * N) if dir_vals is None else np.ascontiguousarray(dir_vals, dtype=np.float64)
    conn = np.ascontiguousarray(mesh.conn, dtype=np.int32)
    flags = np.ascontiguousarray(mesh.flags, dtype=np.uint8)
    assert pos.dtype == np.float64 and pos.flags.c_contiguous and q.dtype == np.float64 and q.flags.c_contiguous
    assert q.size == (d + 1) * N and pos.size == d * N
    stats = np.zeros(5)
    st = L.orc_frac_step(d, N, mesh.n_elems, _i(conn), _u8(flags), _d(pos), _d(q), _d(dir_vals), rho, mu, _d(g), dt,
                         gamma_fs, RES_KIND[residual], min_res, max_iter, rtol, cg_maxit, _d(stats))
    keys = ["picard_iters", "cg_iters_vapp", "cg_iters_pcorr", "cg_iters_v", "final_res"]
    return st, dict(zip(keys, stats.tolist()))


DIRICHLET_T = 32


def boussinesq_step(mesh, pos, q, hc: "HeatCase", alpha, Tr, dir_vals=None, Tbc=None, rho=1000.0, mu=1e-3, g=None,
                    dt=1e-3, solve_heat_first=True, residual="Ax_f", min_res=1e-6, max_iter=10, rtol=1e-10,
                    krylov_maxit=20000, heat_rtol=1e-13):
    """SolverIncompNewton::m_solveBoussinesq without the remesh (IncompNewton/Solver.cpp:249-283):
    the heat equation (on the positions of its turn) and the PSPG step whose F and H factors read
    the node temperatures.  q = (v, p, T), (dim + 2) N, and pos updated in place on success.  The
    heat equation's Dirichlet nodes are those flagged DIRICHLET_T.  Returns (status, which) with
    which "heat" or "momentum" naming the equation that failed."""
    d = mesh.dim
    N = mesh.n_nodes
    nm = (d + 1) * N
    qm = np.ascontiguousarray(q[:nm])
    T = np.ascontiguousarray(q[nm:])
    hflags = np.where(mesh.flags & DIRICHLET_T, mesh.flags | 16, mesh.flags & np.uint8(0xFF ^ 16)).astype(np.uint8)

    def heat():
        hm = type(mesh)(d, pos.copy(), mesh.conn, hflags, getattr(mesh, "h", 1.0))
        st, _, _ = heat_step(hm, T, Tbc=Tbc, hc=hc, dt=dt, rtol=heat_rtol)
        return st

    if solve_heat_first and heat() != 0:
        return 1, "heat"
    with material("Boussinesq", alpha=alpha, Tr=Tr, T=T):
        st, _ = pspg_step(mesh, pos, qm, dir_vals=dir_vals, rho=rho, mu=mu, g=g, dt=dt, residual=residual,
                          min_res=min_res, max_iter=max_iter, rtol=rtol, krylov_maxit=krylov_maxit)
    if st != 0:
        return st, "momentum"
    if not solve_heat_first and heat() != 0:
        return 1, "heat"
    q[:nm] = qm
    q[nm:] = T
    return 0, ""

