/*
 * pfem_gpu.h — C-ABI of the MI355X (gfx950) PFEM assembly + sparse-solve library.
 *
 * The library replaces, per time step, the element-matrix builders, the sparse assembly,
 * the boundary conditions and the sparse solve of ImperatorS79/PFEM's incompressible PSPG
 * solver and the explicit update of its weakly compressible solver.  Everything above
 * those (Lua configuration, the time loop, dt control, CGAL remeshing) stays with the
 * caller.  Reference citations are relative to the reference repository root.
 *
 * Conventions (mirroring the reference's ownership and error behaviour):
 *  - The caller owns every host array and passes pointer + count; the library copies
 *    what it needs and never keeps a host pointer.  The library owns all device memory
 *    behind the opaque handles.  Handles are not copyable.
 *  - Every function returns an int status: 0 OK, < 0 hard error (invalid argument, out of
 *    memory, HIP error; the reference throws std::runtime_error), > 0 numerical failure
 *    (the reference returns false: solver failure, Picard non-convergence, NaN residual).
 *    pfem_last_error() gives the message of the calling thread's last failure.
 *  - A case handle is not re-entrant (one HIP stream per case); distinct case handles may
 *    be driven from distinct host threads.  Indices are int32.
 *  - Vectors are field-blocked like the reference's q vectors (StatesFromToQ.hpp:9-34):
 *    entry (node n, field s) is at n + s*N.  PSPG fields are u, v, [w], p.
 */
#ifndef PFEM_GPU_H
#define PFEM_GPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ---------------------------------------------------------------- */
#define PFEM_OK 0
#define PFEM_E_INVALID (-1)
#define PFEM_E_OOM (-2)
#define PFEM_E_HIP (-3)
#define PFEM_E_UNSUPPORTED (-4)
#define PFEM_E_STATE (-5)
#define PFEM_NOT_CONVERGED 1 /* Krylov iteration limit or Picard iterCount > maxIter */
#define PFEM_NAN 2           /* NaN / breakdown (the reference's NaN residual branch) */

/* ---- node flags (Node.hpp: m_isBound, isFree(), m_isOnFreeSurface, m_isFixed) -------- */
#define PFEM_NODE_BOUND 1
#define PFEM_NODE_FREE 2 /* recomputed by the library: node in no element (Node.inl isFree) */
#define PFEM_NODE_ONFS 4
#define PFEM_NODE_FIXED 8     /* IC.BoundaryFixed: position never updated */
#define PFEM_NODE_DIRICHLET 16 /* boundary node whose tag has a "<Tag>V" BC function (a Conduction case:
                                  "<Tag>T") */
#define PFEM_NODE_DIRICHLET_T 32 /* Boussinesq cases: node whose tag has a "<Tag>T" BC function */

/* ---- enums ------------------------------------------------------------------------- */
#define PFEM_KIND_PSPG 0 /* MomContEqIncompNewton with Solver.id = "PSPG" */
#define PFEM_KIND_WC 1   /* WCompNewtonNoT with Solver.id = "CDS_dpdt" */
#define PFEM_KIND_HEAT 2 /* "Conduction": HeatEqIncompNewton alone, no phase change */
#define PFEM_KIND_FRACSTEP 3 /* MomContEqIncompNewton with Solver.id = "FracStep" */
#define PFEM_KIND_BOUSSINESQ 4 /* "Boussinesq": the PSPG MomContEqIncompNewton with the temperature-dependent
                                  body force, and HeatEqIncompNewton (no phase change), on one mesh */

#define PFEM_RES_U 0    /* MomContEquation.inl:229-236 residual = "U" */
#define PFEM_RES_U_P 1  /* "U_P" */
#define PFEM_RES_AX_F 2 /* "Ax_f" */

#define PFEM_PREC_FP64 0  /* FP64 assembly, SpMV and Krylov */
#define PFEM_PREC_MIXED 1    /* FP64 BiCGSTAB (converges on the FP64 system) preconditioned by
                                the multigrid with FP32 values at level 0 (the residual's
                                product over an FP32 copy of A and the post-smoothing product
                                A P held in FP32; FP64 vectors, coarse levels and coarsest
                                dense inverse), 3D cases from 5,000 nodes; block-Jacobi cases
                                run as FP64.  PFEM_AMG_FP32=1 in the environment when the mesh
                                is set runs a 3D hierarchy wholly in FP32 instead */
#define PFEM_PREC_MIXED_IR 2 /* FP32 inner BiCGSTAB (FP32 vectors) on the FP32 copy of A inside
                                FP64 iterative refinement; falls back to FP64 BiCGSTAB when the
                                FP32 corrections stop reducing the FP64 residual */

#define PFEM_STATE_FP64 0 /* WC node state stored in FP64 */
#define PFEM_STATE_FP32 1 /* WC node state stored in FP32 (accumulators stay FP64) */

typedef struct pfem_ctx pfem_ctx;
typedef struct pfem_case pfem_case;

/* Case parameters.  Replaces the Lua tables read by the equation constructors:
 * MomContEquation.inl:28-30,226-243 (PSPG) and ContEquation.inl:26-44,
 * MomEquation.inl:28-29,149-153 (WC). */
typedef struct pfem_case_desc {
    int32_t dim;  /* 2 or 3 */
    int32_t kind; /* PFEM_KIND_* */
    double rho;   /* Material.rho (PSPG) */
    double mu;    /* Material.mu */
    double gamma; /* Material.gamma (surface tension); > 0 applies the FST facet term of the facets
                     set by pfem_case_set_facets (PSPG, Boussinesq, WC; refused by FracStep) */
    double body_force[3];
    /* WC (ContEq / Material) */
    double K0, K0p, rho_star;
    int32_t meduri;     /* ContEq.stabilization: 1 "Meduri", 0 "None" */
    int32_t wc_state;   /* PFEM_STATE_* */
    /* PSPG Picard (MomContEq table) */
    int32_t residual;   /* PFEM_RES_* */
    int32_t max_iter;   /* maxIter */
    double min_res;     /* minRes */
    /* linear solver (replaces EigenSparseSolver, MomContEquation.hpp:6-12) */
    int32_t precision;  /* PFEM_PREC_* */
    int32_t krylov_max_iter;
    double krylov_rtol; /* stop when ||b - A x|| <= rtol ||b|| (true residual, FP64); inside a
                           Picard step with the Ax_f residual also ||b - A x|| <= 0.1 min_res,
                           so that the inexact solve never decides the absolute Picard test */
    double inner_rtol;  /* PFEM_PREC_MIXED_IR: relative tolerance of each FP32 inner solve */
    int32_t inner_max_iter; /* PFEM_PREC_MIXED_IR: iteration cap of each FP32 inner solve */
    int32_t precond;        /* PFEM_PRECOND_* */
    /* material law of the PSPG momentum equation (MomContEquation.inl:33-66, 100-123) */
    int32_t problem;        /* PFEM_PROBLEM_* */
    double tau0;            /* Bingham yield stress tau0 */
    double m_reg;           /* Bingham regularisation mReg */
    /* heat equation (Material table of HeatEqIncompNewton, HeatEquation.inl:27-32); rho above */
    double cv;              /* cv: specific heat */
    double k_cond;          /* k: conductivity */
    double h_conv;          /* h: convection coefficient of the "<Tag>Qh" facets */
    double t_inf;           /* Tinf: ambient temperature of convection and radiation */
    double eps_rad;         /* epsRad: emissivity of the "<Tag>Qr" facets */
    /* fractional step (MomContEquation.inl:151-161) */
    double gamma_fs;        /* gammaFS in [0, 1]: weight of the previous pressure */
    /* Boussinesq (Material table, MomContEquation.inl:164-197): F factor rho (1 - alpha (T - Tr)),
       PSPG H factor 1 - alpha (T - Tr), T = N T_e at each Gauss point */
    double alpha_exp;       /* alpha: thermal expansion coefficient */
    double t_ref;           /* Tr: reference temperature */
    double dgamma_dt;       /* DgammaDT: the Boussinesq surface-tension factor gamma + DgammaDT (T - Tr) */
} pfem_case_desc;

#define PFEM_PROBLEM_NEWTONIAN 0 /* "IncompNewtonNoT": K factor mu */
#define PFEM_PROBLEM_BINGHAM 1   /* "Bingham": K factor mu + tau0 (1 - exp(-mReg g)) / g, g the element's
                                    strain rate sqrt(V^T B^T ddev B V) in the current node states */

/* BiCGSTAB preconditioner of a PSPG case (chosen at pfem_case_set_mesh). */
#define PFEM_PRECOND_AUTO 0   /* multigrid for 2D cases of >= 2,000 nodes, 3D of >= 50,000 (MIXED: 5,000) */
#define PFEM_PRECOND_JACOBI 1 /* block Jacobi over each node's (dim+1) x (dim+1) diagonal block */
#define PFEM_PRECOND_AMG 2    /* aggregation multigrid V-cycle, block-Jacobi smoothing */

typedef struct pfem_solve_stats {
    int32_t status;       /* PFEM_OK / PFEM_NOT_CONVERGED / PFEM_NAN */
    int32_t krylov_iters; /* BiCGSTAB iterations (inner iterations in mixed mode) */
    int32_t refinements;  /* FP64 refinement sweeps (mixed mode) */
    int32_t restarts;
    double relres;        /* final ||b - A x|| / ||b|| in FP64 */
} pfem_solve_stats;

typedef struct pfem_step_stats {
    int32_t status;
    int32_t picard_iters;
    int32_t krylov_iters;
    int32_t assemblies;
    double residual;   /* final Picard residual */
    double max_relres; /* worst linear-solve relative residual in the step */
    double t_assemble, t_solve, t_update, t_residual; /* seconds (host wall clock) */
} pfem_step_stats;

/* ---- context ------------------------------------------------------------------------
 * One context per device; it owns the stream of pfem_batch_step. */
int pfem_ctx_create(int32_t device, pfem_ctx** out);
int pfem_ctx_destroy(pfem_ctx* ctx);
const char* pfem_last_error(void);
int pfem_version(void);
/* Monotonic clock (nanoseconds) of the serving timestamps. */
int64_t pfem_now_ns(void);

/* ---- case lifetime -------------------------------------------------------------------
 * pfem_case_create replaces the equation constructors (MomContEquation.inl:7-256;
 * MomEquation.inl:8-156 + ContEquation.inl:7-98). */
int pfem_case_create(pfem_ctx* ctx, const pfem_case_desc* desc, pfem_case** out);
int pfem_case_destroy(pfem_case* c);

/* Upload a mesh: the state of Mesh after triangulateAlphaShape / remesh
 * (Mesh.cpp:919-926).  conn: E x (dim+1) node indices (Element::m_nodesIndexes order);
 * flags: N bytes of PFEM_NODE_*; pos: dim x N field-blocked coordinates.  Builds the CSR
 * pattern on the GPU (PSPG) and resets the state to zero. */
int pfem_case_set_mesh(pfem_case* c, int32_t N, int32_t E, const int32_t* conn, const uint8_t* flags,
                       const double* pos);

/* PSPG: q = (dim+1) x N (u, v, [w], p).  WC: v (dim x N), p (N), rho (N), a (dim x N)
 * concatenated, i.e. (2*dim + 2) x N — the node states of WCompNewtonNoT. */
int pfem_case_set_state(pfem_case* c, const double* q);
int pfem_case_get_state(pfem_case* c, double* q);
int pfem_case_set_positions(pfem_case* c, const double* pos);
int pfem_case_get_positions(pfem_case* c, double* pos);
/* Dirichlet velocities, dim x N (read at PFEM_NODE_DIRICHLET nodes): the values the
 * reference obtains from Lua "<Tag>V"(pos, t + dt) (MomContEquationPSPG.inl:208-214,
 * MomEquation.inl:358-364).  Default zero. */
int pfem_case_set_dirichlet(pfem_case* c, const double* vals);
int pfem_case_sync(pfem_case* c);

/* ---- PSPG building blocks -------------------------------------------------------------
 * m_buildAbPSPG + m_applyBCPSPG (MomContEquationPSPG.inl:7-235).  q_prev: previous step
 * state (be and free-node rhs); q_iter: node states used by tau (the current Picard
 * iterate).  Either pointer may be NULL to use the case's device state. */
int pfem_pspg_assemble(pfem_case* c, double dt, const double* q_prev, const double* q_iter);
/* m_solver.analyzePattern/factorize/solve (MomContEquationPSPG.inl:281-290) on the
 * assembled system.  x0 (nullable) is the initial guess; x_out (nullable) receives q. */
int pfem_pspg_solve(pfem_case* c, const double* x0, double* x_out, pfem_solve_stats* stats);
/* CSR export (parity tests).  row_ptr: n+1, col_idx/vals: nnz. */
int pfem_get_csr_sizes(pfem_case* c, int64_t* n, int64_t* nnz);
int pfem_get_csr(pfem_case* c, int32_t* row_ptr, int32_t* col_idx, double* vals);
int pfem_get_rhs(pfem_case* c, double* b);
/* y = A x with the assembled FP64 matrix (host arrays of n). */
int pfem_spmv(pfem_case* c, const double* x, double* y);
/* Instrumentation (no counterpart in the reference): y = A x through the sliced ELL copy of
 * the values that the BiCGSTAB kernels read (the assembly and the BCs write it beside the
 * CSR-order values); for checking the copy against pfem_spmv. */
int pfem_spmv_ell(pfem_case* c, const double* x, double* y);

/* One PSPG time step: PicardAlgo::solve (PicardAlgo.cpp:31-95) with the PSPG callbacks
 * (MomContEquationPSPG.inl:262-372), positions updated as updateNodesPositionFromSave
 * (Mesh.cpp:1238-1277).  On failure (> 0) positions and state are restored
 * (Mesh::restoreNodesList). */
int pfem_pspg_step(pfem_case* c, double dt, pfem_step_stats* stats);
/* pfem_pspg_step, pfem_batch_step and pfem_server_submit take the Dirichlet values once per
 * step, so they return PFEM_E_UNSUPPORTED for a mesh with a boundary Dirichlet node that is
 * not PFEM_NODE_FIXED: its value must be re-evaluated at its new position before every
 * re-assembly (MomContEquationPSPG.inl:206-214), which the callback-level entry points below
 * allow. */

/* ---- the Picard loop's callbacks one by one --------------------------------------------
 * PicardAlgo::solve (PicardAlgo.cpp:31-95) takes three callbacks (NonLinearAlgo.hpp:18-22)
 * that MomContEquationPSPG.inl:262-372 binds.  These entry points let the reference's own
 * PicardAlgo drive the GPU through them, with every vector staying on the device:
 *   prepare  -> pfem_save_positions; pfem_pspg_assemble(c, dt, NULL, NULL)
 *   solve    -> pfem_pspg_solve(c, NULL, NULL, &st); pfem_update_positions(c, dt, NULL);
 *               [pfem_case_set_dirichlet with the BC values at the new positions;]
 *               pfem_pspg_assemble(c, dt, NULL, NULL)   (on failure: pfem_restore_positions)
 *   residual -> pfem_residual(c, kind, NULL, NULL, &res)
 * and, when PicardAlgo::solve returns true, pfem_pspg_commit (the next step's qPrev is read
 * from the node states, MomContEquation.inl:287-298).  The device keeps the case's node
 * states as the Picard iterate q (what tau reads and what pfem_pspg_solve writes) and the
 * step's previous state q_prev. */
/* Mesh::saveNodesList (Mesh.cpp:1047-1053) at the start of a step: saves the positions; the
 * node states and the previous iterate are set to q_prev (PicardAlgo.cpp:37-42). */
int pfem_save_positions(pfem_case* c);
/* Mesh::restoreNodesList (Mesh.cpp:1019-1045): positions back to the saved ones, node states
 * back to q_prev. */
int pfem_restore_positions(pfem_case* c);
/* setNodesStatesfromQ + Mesh::updateNodesPositionFromSave (MomContEquationPSPG.inl:293-295,
 * Mesh.cpp:1238-1277): node states = q (host array of (dim+1) N, or NULL for the device
 * iterate the last pfem_pspg_solve wrote), then x = x_saved + dt q_v for non-fixed nodes. */
int pfem_update_positions(pfem_case* c, double dt, const double* q);
/* Residual callback (MomContEquationPSPG.inl:314-372) of kind PFEM_RES_*: U / U_P from the
 * iterate and the previous iterate, Ax_f = ||A q - b|| with the system last assembled.
 * q_iter / q_iter_prev: host arrays, or NULL for the device copies.  Afterwards the device's
 * previous iterate is the current one (PicardAlgo.cpp:72-73). */
int pfem_residual(pfem_case* c, int32_t kind, const double* q_iter, const double* q_iter_prev, double* out);
/* End of a converged step: q_prev = the node states (MomContEquation.inl:287-298 reads the
 * next step's qPrev from them). */
int pfem_pspg_commit(pfem_case* c);
/* Copy of the device iterate (the node states) into q ((dim+1) N). */
int pfem_get_iterate(pfem_case* c, double* q);

/* ---- WC ------------------------------------------------------------------------------
 * One CDS_dpdt step (WCompNewton/Solver.cpp:236-276). */
int pfem_wc_step(pfem_case* c, double dt, pfem_step_stats* stats);
/* computeNextDT (WCompNewton/Solver.cpp:192-234): dt = min(sqrt(min_e sc^2 he^2 / S), max_dt). */
int pfem_wc_next_dt(pfem_case* c, double security_coeff, double max_dt, double* dt_out);

/* ---- extractors ------------------------------------------------------------------------
 * The problem's global "mass", as MassExtractor::update (extractors/MassExtractor.cpp:27-38)
 * reads it: PSPG, the sum of the element sizes (IncompNewton/Problem.cpp:173-190); WC, each
 * size times the mean density of the element's nodes (WCompNewton/Problem.cpp:216-240). */
int pfem_case_mass(pfem_case* c, double* mass);
/* MinMaxExtractor::update (extractors/MinMaxExtractor.cpp:28-62): the smallest (want_max = 0)
 * or largest coordinate `coordinate` (0 .. dim-1) over the nodes that are neither bound nor
 * free; the largest starts from 0, as the reference's does. */
int pfem_case_minmax(pfem_case* c, int32_t coordinate, int32_t want_max, double* value);
/* No counterpart in the reference: pfem_case_mass for callers that cannot allocate or
 * synchronise (HIP graph capture, tight serving loops).  The same kernels, with the
 * per-workgroup partial sums in a fixed 16 KiB device array of the library (2,048 doubles per
 * GPU, held until the process ends; the element loop is grid-strided over at most 2,048
 * workgroups, so every mesh fits) instead of the case's buffer, and the sum written to the
 * device double d_mass (added in the same order as pfem_case_mass, so the same bits).
 * Enqueued on the case's stream without waiting: read d_mass after pfem_case_sync or a
 * later operation on that stream.  Allocates nothing.  Every call on one GPU shares the
 * array: calls that overlap on one GPU (cases on different streams at the same time)
 * race. */
int pfem_case_mass_noalloc(pfem_case* c, double* d_mass);

/* ---- batched serving (no counterpart in the reference: many independent cases) -------
 * Runs one time step of every listed case as one batch: each operation of the cases'
 * steps (assembly, BC, BiCGSTAB iterations, residuals, WC passes) is one launch over a
 * device work table of all cases that need it.  Cases must be distinct handles on the
 * same context.  status[i] receives each case's step status; stats may be NULL.  Each
 * case's result is bitwise the same as stepping it alone (PSPG). */
int pfem_batch_step(pfem_ctx* ctx, int32_t ncases, pfem_case* const* cases, const double* dts, int32_t* status,
                    pfem_step_stats* stats);

/* Continuous-batching server: a request is one time step of one case (a PSPG Picard step,
 * or a WC step followed by computeNextDT).  Requests are admitted in arrival order when an
 * engine slot is free and the case's step workspace (matrix values, rhs, preconditioner,
 * Picard iterates, Krylov vectors: pfem_case_workspace_bytes) fits in the server's HBM
 * arena; the mesh, pattern and state stay resident with the case.  A request that
 * arrives while others run joins the running batch at the next engine tick.  The server
 * runs on its own host thread and HIP stream; a case with a pending request must not be
 * used by any other call until its result has been polled. */
typedef struct pfem_server pfem_server;

typedef struct pfem_server_desc {
    int64_t hbm_budget;       /* bytes of the workspace arena; 0: half of the free device memory */
    int32_t max_inflight;     /* requests in flight at once (engine slots), 1..4096 */
    int32_t max_chunk;        /* BiCGSTAB iterations per engine tick at most (0: 64) */
    double wc_security_coeff; /* CFL securityCoeff of the WC requests' next time step */
    double wc_max_dt;         /* maxDT of the WC requests' next time step */
} pfem_server_desc;

typedef struct pfem_server_result {
    int64_t tag;
    int32_t status; /* the step's status (PFEM_OK, PFEM_NOT_CONVERGED, PFEM_NAN, < 0 errors) */
    int32_t kind;   /* PFEM_KIND_* */
    pfem_step_stats stats;
    double next_dt; /* WC: computeNextDT after the step */
    int64_t t_arrival_ns, t_admit_ns, t_done_ns; /* pfem_now_ns clock */
} pfem_server_result;

int pfem_server_create(pfem_ctx* ctx, const pfem_server_desc* desc, pfem_server** out);
/* Waits for every submitted request, then frees the server (unpolled results are dropped). */
int pfem_server_destroy(pfem_server* s);
/* Queues one step of `c` with time step dt.  arrival_ns (pfem_now_ns clock; <= 0: now) is
 * the request's arrival: it is not admitted before then and its latency counts from it.
 * PFEM_E_OOM when the case's workspace exceeds the arena, PFEM_E_STATE when the case
 * already has a pending request. */
int pfem_server_submit(pfem_server* s, pfem_case* c, double dt, int64_t tag, int64_t arrival_ns);
/* Moves up to max completed results into out (*n_out of them), waiting up to timeout_ms
 * (< 0: until one is available) when none is ready. */
int pfem_server_poll(pfem_server* s, pfem_server_result* out, int32_t max, int32_t timeout_ms, int32_t* n_out);
/* {engine ticks, peak requests in flight, arena bytes, arena peak bytes, submitted,
 *  completed, arena bytes in use, requests not yet arrived} */
int pfem_server_stats(pfem_server* s, int64_t* out8);

/* Instrumentation (no counterpart in the reference): times, with HIP events on the case's
 * stream, the first SpMV launch of every BiCGSTAB iteration of single-case solves.
 * out4 (nullable) receives the statistics gathered since the last call:
 * {avg ms FP64, launches FP64, avg ms FP32, launches FP32}; enable = 1 then resets and
 * (re)starts timing, enable = 0 stops it. */
int pfem_profile_spmv(pfem_case* c, int32_t enable, double* out4);

/* Instrumentation (no counterpart in the reference): back-to-back timings, in
 * microseconds per launch, of kernels on this case's assembled system.  out10: {the FP64
 * BiCGSTAB SpMV kernel (v = A ph with the fused (rh, v)) as a solve launches it, node-block
 * SpMV FP32 (0 unless PFEM_PREC_MIXED_IR), streaming read of the nnz FP64 values, copy
 * kernel, bytes moved by the copy, nnz, node-graph nnz, lanes per node, the multigrid's
 * level-0 residual t = r - A w over the FP64 ELL copy as a V-cycle launches it (0 without a
 * multigrid), its workgroup count}.  The case needs an assembled system and its step
 * workspace. */
int pfem_bench_kernels(pfem_case* c, int32_t reps, double* out10);
/* Instrumentation: XCD (XCC) id on which each of nblocks single-wavefront workgroups ran. */
int pfem_probe_xcd(pfem_ctx* ctx, int32_t nblocks, int32_t* out);

/* Device footprint of a case in bytes: the resident mesh, pattern and state plus the
 * step workspace. */
int pfem_case_footprint(pfem_case* c, int64_t* bytes);
/* Instrumentation: host seconds of the last pfem_case_set_mesh (the work a remesh triggers)
 * and what it found: {PSPG pattern build (node graph, pair-element incidence, CSR rows),
 * sliced ELL structure, multigrid hierarchy (host aggregation + device set-up), total
 * including uploads, smallest element shape quality 6 sqrt(2) |e| / l_rms^3 (3D; 1 in 2D),
 * the multigrid smoother damping it implies (0.9, or 0.5 below quality 0.1)}. */
int pfem_case_setup_times(pfem_case* c, double* out6);
/* Diagnostics of the case's own engine (the direct calls; no counterpart in the reference):
 * {ticks, replays of captured BiCGSTAB iteration chunks, captures, peak requests in flight}.
 * Chunks are captured in HIP graphs when the environment sets PFEM_GRAPH=1. */
int pfem_case_engine_stats(pfem_case* c, int64_t* out4);
/* Bytes of the step workspace alone (what a serving request borrows from the arena). */
int pfem_case_workspace_bytes(pfem_case* c, int64_t* bytes);
/* Allocates the case's own step workspace and direct-call resources now (the direct calls
 * otherwise allocate them at their first use), or frees the workspace (an idle case keeps only its mesh, pattern and state; a
 * served case then borrows the server's arena again).  The assembled system and the Picard
 * iterate are lost on release. */
int pfem_case_reserve_workspace(pfem_case* c);
int pfem_case_release_workspace(pfem_case* c);

/* ---- Heat (conduction) -----------------------------------------------------------------------
 * HeatEqIncompNewton without phase change (srcs/simulation/physics/IncompNewton/HeatEquation.inl),
 * as SolverIncompNewton::m_solveConduction runs it (IncompNewton/Solver.cpp:286-296): one node
 * state T, nodes fixed.  pfem_case_set_state / get_state take N temperatures; the node flag
 * PFEM_NODE_DIRICHLET marks the nodes whose tag has a "<Tag>T" function. */
#define PFEM_HEAT_Q 1  /* facet with a prescribed flux "<Tag>Q" (every node's tag has it) */
#define PFEM_HEAT_QH 2 /* facet with convection h (T - Tinf) ("<Tag>Qh") */
#define PFEM_HEAT_QR 4 /* facet with radiation sigma epsRad (T^4 - Tinf^4) ("<Tag>Qr") */
/* The boundary facets (pfem_mesher_get_facets: nodes F x dim, out node F) and each one's BC
 * kinds, PFEM_HEAT_* bits: m_applyBC (HeatEquation.inl:289-348) applies a kind to a facet whose
 * nodes all carry it, with the tag "FreeSurface" for a facet on the free surface. */
int pfem_heat_set_facets(pfem_case* c, int32_t F, const int32_t* nodes, const int32_t* out_node,
                         const uint8_t* kinds);
/* Gauss points of every facet (F x 3 x dim; Facet::getPosFromGP), where the caller evaluates
 * the flux "<Tag>Q"(pos, t + dt) of PFEM_HEAT_Q facets. */
int pfem_heat_facet_points(pfem_case* c, double* xgp);
/* The flux vectors at those points (F x 3 x dim), or NULL for none. */
int pfem_heat_set_flux(pfem_case* c, const double* q);
/* Dirichlet temperatures, N values ("<Tag>T"(pos, t + dt), HeatEquation.inl:352-356). */
int pfem_heat_set_dirichlet(pfem_case* c, const double* t_bc);
/* m_buildAb + m_applyBC (HeatEquation.inl:209-364) from the node temperatures, for checks:
 * afterwards pfem_heat_get_system exports the system. */
int pfem_heat_assemble(pfem_case* c, double dt);
int pfem_heat_system_nnz(pfem_case* c, int64_t* nnz);
/* CSR rows N + 1, columns and values nnz, rhs N. */
int pfem_heat_get_system(pfem_case* c, int32_t* row_ptr, int32_t* col_idx, double* vals, double* b);
/* One HeatEqIncompNewton::solve: the system, the conjugate-gradient solve with the diagonal
 * preconditioner from the current temperatures (the reference's Eigen::ConjugateGradient,
 * HeatEquation.hpp:50) to krylov_rtol, the new temperatures.  PFEM_NOT_CONVERGED / PFEM_NAN
 * leave the temperatures unchanged. */
int pfem_heat_step(pfem_case* c, double dt, pfem_step_stats* stats);

/* ---- Surface tension ---------------------------------------------------------------------------
 * The boundary facets of the case's mesh (Mesh::m_facets: nodes F x dim in Facet::m_nodesIndexes
 * order, and each facet's element; pfem_mesher_get_facets gives both) for the FST term of
 * getFST (MatricesBuilder.inl:389-402), which m_applyBCPSPG (MomContEquationPSPG.inl:155-186)
 * adds to the momentum rows of the facet's element for every facet with a node on the free
 * surface (PFEM_NODE_ONFS), and the WC m_applyBC (WCompNewton/MomEquation.inl:312-335) to the
 * momentum force for every facet whose nodes all are; the factor is gamma, or for a Boussinesq
 * case gamma + DgammaDT (T - Tr) at the facet's Gauss points.  pfem_case_set_mesh clears the
 * facets; pfem_case_set_mesh_from / pfem_case_remesh set the mesher's facets when the case has
 * surface tension.  F = 0 clears them. */
int pfem_case_set_facets(pfem_case* c, int32_t F, const int32_t* nodes, const int32_t* elem);

/* ---- Fractional step (PFEM_KIND_FRACSTEP) --------------------------------------------------
 * One MomContEqIncompNewton::solve with Solver.id = "FracStep" (MomContEquation.inl:296-303;
 * the Picard callbacks of MomContEquationFracStep.inl:462-630): per Picard iteration the
 * velocity prediction (M/dt + K) vt = F + M/dt v_prev + gammaFS D^T p_prev (free rows
 * v_prev + dt g, Dirichlet rows and columns from pfem_case_set_dirichlet), the pressure
 * correction L p = -(rho/dt) D vt + gammaFS L p_prev (zero right-hand side on free-surface and
 * free nodes), the velocity correction M dv = dt D^T (p - gammaFS p_prev), v = vt + dv,
 * positions = saved + dt v; until the U / U_P residual (residual, min_res, max_iter of the
 * descriptor; "Ax_f" is refused at pfem_case_create, as the reference throws) drops below
 * min_res.  Each system is solved by diagonal-preconditioned conjugate gradient from a zero
 * guess (the reference's Eigen::ConjugateGradient<Lower|Upper>) to krylov_rtol within
 * krylov_max_iter iterations.  The state is q = (v, p), (dim + 1) N values, as for PSPG.
 * PFEM_NOT_CONVERGED (Picard iterations or a CG solve) / PFEM_NAN restore the positions and
 * leave the state unchanged.  Dirichlet nodes must be fixed (as for pfem_pspg_step). */
int pfem_frac_step(pfem_case* c, double dt, pfem_step_stats* stats);

/* ---- Host remesher (no GPU work) ------------------------------------------------------------
 * The reference's Mesh class after loadFromFile: triangulateAlphaShape (Mesh2D.cpp:32-251,
 * Mesh3D.cpp:36-286) and remesh (Mesh.cpp:919-926 = laplacianSmoothingBoundaries, addNodes,
 * removeNodes, checkBoundingBox, triangulateAlphaShape), without CGAL: an exact-predicate
 * Delaunay triangulation with the lexicographic symbolic perturbation, whose element set is the
 * one CGAL's builds.  Elements come out in a canonical order (sorted node tuples, smallest node
 * first, positive orientation), not CGAL's iteration order; the nodes addNodes appends are
 * numbered in that order.  One mesher per case; distinct meshers may run on distinct host
 * threads concurrently. */
typedef struct pfem_mesher pfem_mesher;

typedef struct pfem_mesher_desc {
    int32_t dim;
    double hchar, alpha, omega, gamma;   /* MeshCreateInfo (Mesh.hpp:19-32): Mesh.hchar, alpha, omega, gamma */
    int32_t add_on_fs;                   /* Mesh.addOnFS */
    int32_t delete_flying_nodes;         /* Mesh.deleteFlyingNodes */
    int32_t laplacian_smoothing_boundaries; /* Mesh.laplacianSmoothingBoundaries */
    int32_t has_bounding_box;            /* 0: no node is removed for leaving a box */
    double bounding_box[6];              /* Mesh.boundingBox: [xmin, ymin, (zmin,) xmax, ymax, (zmax)] */
    int32_t n_exclusion_zones;           /* Mesh.exclusionZones: elements whose centre lies in one are dropped */
    const double* exclusion_zones;       /* n_exclusion_zones x 2 dim, bounding-box layout (copied) */
} pfem_mesher_desc;

/* Mesh::Mesh (Mesh.cpp:12-25) without the file read (gmsh_reader.py builds the node cloud). */
int pfem_mesher_create(const pfem_mesher_desc* desc, pfem_mesher** out);
int pfem_mesher_destroy(pfem_mesher* m);
/* The node cloud: pos dim x n, states nstates x n (field-blocked, nullable), flags n bytes of
 * PFEM_NODE_* (BOUND, FIXED, DIRICHLET are kept; FREE and ONFS are computed), tags n (nullable;
 * Node::m_tag, -1 for nodes remeshing adds).  Clears the triangulation. */
int pfem_mesher_set_nodes(pfem_mesher* m, int32_t n, int32_t nstates, const double* pos, const double* states,
                          const uint8_t* flags, const int32_t* tags);
/* Positions and node states after time steps, same node count (either pointer nullable). */
int pfem_mesher_set_positions_states(pfem_mesher* m, const double* pos, const double* states);
/* Mesh::triangulateAlphaShape (Mesh.cpp:1055-1061). */
int pfem_mesher_triangulate(pfem_mesher* m);
/* Mesh::remesh (Mesh.cpp:919-926); needs a triangulation. */
int pfem_mesher_remesh(pfem_mesher* m);
int pfem_mesher_sizes(pfem_mesher* m, int32_t* n_nodes, int32_t* n_elems, int32_t* n_facets);
/* pos dim x N, states nstates x N, flags N, tags N: any may be null. */
int pfem_mesher_get_nodes(pfem_mesher* m, double* pos, double* states, uint8_t* flags, int32_t* tags);
/* E x (dim+1) node indices, detJ > 0. */
int pfem_mesher_get_elements(pfem_mesher* m, int32_t* conn);
/* Free-surface facets (Mesh2D.cpp:192-242, Mesh3D.cpp:215-282): nodes F x dim (oriented so that
 * the out node, the element's vertex opposite the facet, is on the positive side), out_node F,
 * elem F; any may be null. */
int pfem_mesher_get_facets(pfem_mesher* m, int32_t* nodes, int32_t* out_node, int32_t* elem);
/* Last triangulate / remesh: {nodes added, removed, out of the box, flying nodes deleted,
 * duplicate positions not inserted, seconds in node operations, in the Delaunay triangulation,
 * in the alpha shape and adjacency}. */
int pfem_mesher_stats(pfem_mesher* m, double* out8);
/* Remeshes a case between steps: reads its positions and node states (PSPG (dim+1) states per
 * node, WC 2 dim + 2), hands them to the mesher (which holds the case's last triangulation),
 * remeshes, and sets the new mesh and states on the case (pfem_case_set_mesh,
 * pfem_case_set_state).  Dirichlet values are per node: set them again afterwards.  Replaces
 * the m_pMesh->remesh call of SolverIncompNewton::m_solveIncompNewtonNoT
 * (IncompNewton/Solver.cpp:236-246) and SolverWCompNewton (WCompNewton/Solver.cpp:262-270). */
int pfem_case_remesh(pfem_case* c, pfem_mesher* m);
/* Sets the mesher's current mesh and node states on a case (after pfem_mesher_triangulate). */
int pfem_case_set_mesh_from(pfem_case* c, pfem_mesher* m);

#ifdef __cplusplus
}
#endif

#endif /* PFEM_GPU_H */
