"""ctypes binding of libpfem_gpu.so (include/pfem_gpu.h).

The HIP library is the only implementation: if it is not built, loading fails with an
error instead of falling back to anything else.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libpfem_gpu.so")

PFEM_OK = 0
PFEM_E_INVALID = -1
PFEM_E_OOM = -2
PFEM_E_HIP = -3
PFEM_E_UNSUPPORTED = -4
PFEM_E_STATE = -5
PFEM_NOT_CONVERGED = 1
PFEM_NAN = 2

KIND_PSPG = 0
KIND_WC = 1
KIND_HEAT = 2
KIND_FRACSTEP = 3
KIND_BOUSSINESQ = 4
HEAT_Q, HEAT_QH, HEAT_QR = 1, 2, 4
RES = {"U": 0, "U_P": 1, "Ax_f": 2}
PREC_FP64 = 0
PREC_MIXED = 1
PREC_MIXED_IR = 2
PRECISION = {"fp64": PREC_FP64, "mixed": PREC_MIXED, "mixed_ir": PREC_MIXED_IR}
PROBLEM = {"Newtonian": 0, "Bingham": 1}
PRECOND = {"auto": 0, "jacobi": 1, "amg": 2}
STATE_FP64 = 0
STATE_FP32 = 1

EXPORTS = [
    "pfem_ctx_create", "pfem_ctx_destroy", "pfem_last_error", "pfem_version", "pfem_case_create",
    "pfem_case_destroy", "pfem_case_set_mesh", "pfem_case_set_state", "pfem_case_get_state",
    "pfem_case_set_positions", "pfem_case_get_positions", "pfem_case_set_dirichlet", "pfem_case_sync",
    "pfem_pspg_assemble", "pfem_pspg_solve", "pfem_get_csr_sizes", "pfem_get_csr", "pfem_get_rhs", "pfem_spmv",
    "pfem_pspg_step", "pfem_wc_step", "pfem_wc_next_dt", "pfem_case_mass", "pfem_case_minmax", "pfem_batch_step",
    "pfem_case_footprint",
    "pfem_profile_spmv", "pfem_bench_kernels", "pfem_probe_xcd", "pfem_now_ns", "pfem_case_workspace_bytes",
    "pfem_server_create", "pfem_server_destroy", "pfem_server_submit", "pfem_server_poll", "pfem_server_stats",
    "pfem_case_reserve_workspace", "pfem_case_release_workspace", "pfem_save_positions", "pfem_restore_positions",
    "pfem_update_positions", "pfem_residual", "pfem_pspg_commit", "pfem_get_iterate", "pfem_case_mass_noalloc",
    "pfem_spmv_ell", "pfem_mesher_create", "pfem_mesher_destroy", "pfem_mesher_set_nodes",
    "pfem_mesher_set_positions_states", "pfem_mesher_triangulate", "pfem_mesher_remesh", "pfem_mesher_sizes",
    "pfem_mesher_get_nodes", "pfem_mesher_get_elements", "pfem_mesher_get_facets", "pfem_mesher_stats",
    "pfem_case_remesh", "pfem_case_set_mesh_from", "pfem_case_setup_times", "pfem_heat_set_facets",
    "pfem_heat_facet_points", "pfem_heat_set_flux", "pfem_heat_set_dirichlet", "pfem_heat_assemble",
    "pfem_heat_system_nnz", "pfem_heat_get_system", "pfem_heat_step", "pfem_frac_step", "pfem_case_set_facets", "pfem_case_engine_stats",
]


class CaseDesc(C.Structure):
    _fields_ = [
        ("dim", C.c_int32), ("kind", C.c_int32), ("rho", C.c_double), ("mu", C.c_double), ("gamma", C.c_double),
        ("body_force", C.c_double * 3), ("K0", C.c_double), ("K0p", C.c_double), ("rho_star", C.c_double),
        ("meduri", C.c_int32), ("wc_state", C.c_int32), ("residual", C.c_int32), ("max_iter", C.c_int32),
        ("min_res", C.c_double), ("precision", C.c_int32), ("krylov_max_iter", C.c_int32),
        ("krylov_rtol", C.c_double), ("inner_rtol", C.c_double), ("inner_max_iter", C.c_int32), ("precond", C.c_int32),
        ("problem", C.c_int32), ("tau0", C.c_double), ("m_reg", C.c_double),
        ("cv", C.c_double), ("k_cond", C.c_double), ("h_conv", C.c_double), ("t_inf", C.c_double),
        ("eps_rad", C.c_double),
        ("gamma_fs", C.c_double),
        ("alpha_exp", C.c_double),
        ("t_ref", C.c_double),
        ("dgamma_dt", C.c_double),
    ]


class MesherDesc(C.Structure):
    _fields_ = [
        ("dim", C.c_int32), ("hchar", C.c_double), ("alpha", C.c_double), ("omega", C.c_double),
        ("gamma", C.c_double), ("add_on_fs", C.c_int32), ("delete_flying_nodes", C.c_int32),
        ("laplacian_smoothing_boundaries", C.c_int32), ("has_bounding_box", C.c_int32),
        ("bounding_box", C.c_double * 6), ("n_exclusion_zones", C.c_int32), ("exclusion_zones", C.POINTER(C.c_double)),
    ]


class SolveStats(C.Structure):
    _fields_ = [("status", C.c_int32), ("krylov_iters", C.c_int32), ("refinements", C.c_int32),
                ("restarts", C.c_int32), ("relres", C.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class StepStats(C.Structure):
    _fields_ = [("status", C.c_int32), ("picard_iters", C.c_int32), ("krylov_iters", C.c_int32),
                ("assemblies", C.c_int32), ("residual", C.c_double), ("max_relres", C.c_double),
                ("t_assemble", C.c_double), ("t_solve", C.c_double), ("t_update", C.c_double),
                ("t_residual", C.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class ServerDesc(C.Structure):
    _fields_ = [("hbm_budget", C.c_int64), ("max_inflight", C.c_int32), ("max_chunk", C.c_int32),
                ("wc_security_coeff", C.c_double), ("wc_max_dt", C.c_double)]


class ServerResult(C.Structure):
    _fields_ = [("tag", C.c_int64), ("status", C.c_int32), ("kind", C.c_int32), ("stats", StepStats),
                ("next_dt", C.c_double), ("t_arrival_ns", C.c_int64), ("t_admit_ns", C.c_int64),
                ("t_done_ns", C.c_int64)]

    def as_dict(self):
        d = {k: getattr(self, k) for k, _ in self._fields_ if k != "stats"}
        d["stats"] = self.stats.as_dict()
        return d


_lib = None
_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int32)
_u8p = C.POINTER(C.c_uint8)
_vp = C.c_void_p


def lib():
    """Load libpfem_gpu.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m pfem_amd.build` "
                           "(the HIP library is required; there is no fallback)")
    L = C.CDLL(LIB_PATH)
    sig = {
        "pfem_ctx_create": [C.c_int32, C.POINTER(_vp)],
        "pfem_ctx_destroy": [_vp],
        "pfem_last_error": [],
        "pfem_version": [],
        "pfem_case_create": [_vp, C.POINTER(CaseDesc), C.POINTER(_vp)],
        "pfem_case_destroy": [_vp],
        "pfem_case_set_mesh": [_vp, C.c_int32, C.c_int32, _ip, _u8p, _dp],
        "pfem_case_set_state": [_vp, _dp],
        "pfem_case_get_state": [_vp, _dp],
        "pfem_case_set_positions": [_vp, _dp],
        "pfem_case_get_positions": [_vp, _dp],
        "pfem_case_set_dirichlet": [_vp, _dp],
        "pfem_case_sync": [_vp],
        "pfem_pspg_assemble": [_vp, C.c_double, _dp, _dp],
        "pfem_pspg_solve": [_vp, _dp, _dp, C.POINTER(SolveStats)],
        "pfem_get_csr_sizes": [_vp, C.POINTER(C.c_int64), C.POINTER(C.c_int64)],
        "pfem_get_csr": [_vp, _ip, _ip, _dp],
        "pfem_get_rhs": [_vp, _dp],
        "pfem_spmv": [_vp, _dp, _dp],
        "pfem_pspg_step": [_vp, C.c_double, C.POINTER(StepStats)],
        "pfem_wc_step": [_vp, C.c_double, C.POINTER(StepStats)],
        "pfem_wc_next_dt": [_vp, C.c_double, C.c_double, _dp],
        "pfem_case_mass": [_vp, _dp],
        "pfem_case_minmax": [_vp, C.c_int32, C.c_int32, _dp],
        "pfem_batch_step": [_vp, C.c_int32, C.POINTER(_vp), _dp, _ip, C.POINTER(StepStats)],
        "pfem_case_footprint": [_vp, C.POINTER(C.c_int64)],
        "pfem_profile_spmv": [_vp, C.c_int32, _dp],
        "pfem_bench_kernels": [_vp, C.c_int32, _dp],
        "pfem_probe_xcd": [_vp, C.c_int32, _ip],
        "pfem_now_ns": [],
        "pfem_case_workspace_bytes": [_vp, C.POINTER(C.c_int64)],
        "pfem_server_create": [_vp, C.POINTER(ServerDesc), C.POINTER(_vp)],
        "pfem_server_destroy": [_vp],
        "pfem_server_submit": [_vp, _vp, C.c_double, C.c_int64, C.c_int64],
        "pfem_server_poll": [_vp, C.POINTER(ServerResult), C.c_int32, C.c_int32, _ip],
        "pfem_server_stats": [_vp, C.POINTER(C.c_int64)],
        "pfem_case_reserve_workspace": [_vp],
        "pfem_case_release_workspace": [_vp],
        "pfem_save_positions": [_vp],
        "pfem_restore_positions": [_vp],
        "pfem_update_positions": [_vp, C.c_double, _dp],
        "pfem_residual": [_vp, C.c_int32, _dp, _dp, _dp],
        "pfem_pspg_commit": [_vp],
        "pfem_get_iterate": [_vp, _dp],
        "pfem_case_mass_noalloc": [_vp, _vp],
        "pfem_spmv_ell": [_vp, _dp, _dp],
        "pfem_mesher_create": [C.POINTER(MesherDesc), C.POINTER(_vp)],
        "pfem_mesher_destroy": [_vp],
        "pfem_mesher_set_nodes": [_vp, C.c_int32, C.c_int32, _dp, _dp, _u8p, _ip],
        "pfem_mesher_set_positions_states": [_vp, _dp, _dp],
        "pfem_mesher_triangulate": [_vp],
        "pfem_mesher_remesh": [_vp],
        "pfem_mesher_sizes": [_vp, _ip, _ip, _ip],
        "pfem_mesher_get_nodes": [_vp, _dp, _dp, _u8p, _ip],
        "pfem_mesher_get_elements": [_vp, _ip],
        "pfem_mesher_get_facets": [_vp, _ip, _ip, _ip],
        "pfem_mesher_stats": [_vp, _dp],
        "pfem_case_remesh": [_vp, _vp],
        "pfem_case_set_mesh_from": [_vp, _vp],
        "pfem_case_setup_times": [_vp, _dp],
        "pfem_heat_set_facets": [_vp, C.c_int32, _ip, _ip, _u8p],
        "pfem_heat_facet_points": [_vp, _dp],
        "pfem_heat_set_flux": [_vp, _dp],
        "pfem_heat_set_dirichlet": [_vp, _dp],
        "pfem_heat_assemble": [_vp, C.c_double],
        "pfem_heat_system_nnz": [_vp, C.POINTER(C.c_int64)],
        "pfem_heat_get_system": [_vp, _ip, _ip, _dp, _dp],
        "pfem_heat_step": [_vp, C.c_double, C.POINTER(StepStats)],
        "pfem_frac_step": [_vp, C.c_double, C.POINTER(StepStats)],
        "pfem_case_set_facets": [_vp, C.c_int32, _ip, _ip],
        "pfem_case_engine_stats": [_vp, C.POINTER(C.c_int64)],
    }
    for name, args in sig.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = C.c_int
    L.pfem_last_error.restype = C.c_char_p
    L.pfem_now_ns.restype = C.c_int64
    _lib = L
    return L


class PfemError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"[{code}] {msg}")
        self.code = code


def check(rc: int, allow_positive: bool = False) -> int:
    """< 0 raises (the reference throws); > 0 is a numerical failure (the reference returns false)."""
    if rc < 0 or (rc > 0 and not allow_positive):
        msg = lib().pfem_last_error()
        raise PfemError(rc, msg.decode() if msg else "")
    return rc
