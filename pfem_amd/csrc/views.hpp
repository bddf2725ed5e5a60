// Plain-old-data views of one case's device arrays, passed to kernels through device
// arrays indexed by BlockTask::case_id.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

namespace pfem {

// Surface-tension facets of a case (fst.hpp gathers them per node).
struct FstView {
    int F = 0;                  // facets (0: no surface tension)
    int all_nodes = 0;          // 1: a facet acts when every node is on the free surface (WC), 0: any node
    const int* nodes = nullptr; // [F][dim] facet nodes (Facet::m_nodesIndexes order)
    const int* elem = nullptr;  // [F] element of each facet
    const int* ptr = nullptr;   // [N+1] node -> entries
    const unsigned* ent = nullptr;  // f << 2 | a, a the node's local index in the facet's element
    const double* T = nullptr;  // Boussinesq node temperatures (null: the factor is gamma)
    double gamma = 0, dgdT = 0, t_ref = 0;
};

// PSPG assembly + boundary conditions (MomContEquationPSPG.inl:7-235).
struct PspgView {
    int dim, N, E, mom_nnz;
    const int* conn;          // [ne][E]
    const double* pos;        // [dim][N]
    const double* q_prev;     // [(dim+1)N]  states at the start of the time step (be, free-node rhs)
    const double* q_iter;     // [(dim+1)N]  current Picard iterate (node states used by tau)
    const int4* node_info;    // [N] {mom_ptr, cont_ptr, deg, flags}
    const int* nbr_ptr;       // [N+1] node graph
    const int* nbr_idx;       // [graph_nnz]
    const int* pei_ptr;       // [graph_nnz+1] pair-element incidence
    const uint32_t* pei;      // [E*ne*ne]  e << 4 | a << 2 | b, sorted by e within each pair
    const double* dir_vals;   // [dim][N] Dirichlet velocities (read at Dirichlet nodes)
    double* vals;             // [nnz]
    double* b;                // [(dim+1)N]
    float* vals32;            // [nnz] FP32 copy for the mixed-precision inner solve (may be null)
    const int* diag_pos;      // [n]
    int ell_nsl, pad_ell;     // ELL copy of the values (EllMat): slices,
    const int* ell_soff;      // slice offsets,
    double* vals_ell;         // values written after assembly and BC (null: none)
    float* vals32_ell;        // FP32 ELL copy for an FP32 multigrid hierarchy (null: none)
    double* erec;             // [E][16] element records of the assembly (pspg.hip)
    double* esv;              // [E][4] sums of the previous-step velocities over each element
                              //        (and, with T, of the node temperatures in the 4th slot)
    double* binv;             // [(dim+1)^2][N] inverse diagonal node blocks (block Jacobi), SoA
    float* binv32;            // FP32 copy for the mixed-precision inner solve (may be null)
    long long nnz;
    double rho, mu, dt;
    double g[3];
    int problem, pad_problem;  // PFEM_PROBLEM_*
    double tau0, mreg;         // Bingham parameters
    // Boussinesq: node temperatures (null: none) scaling the body force of F by
    // 1 - alpha (T - Tr) at each Gauss point, and H by the same factor
    const double* T;
    double alpha, t_ref;
    FstView fst;  // surface tension (fst.F == 0: none)
};

// The assembled PSPG matrix seen through its node structure.  Values are the CSR values
// (pattern.hpp layout); the column structure is read from the node graph instead of the
// CSR column indices: the full rows (i, s) of node i all hold, for t = 0..nf-1, the
// columns nbr_k + t*N of i's deg neighbours at offset t*deg + k, and every other row
// holds only its diagonal.
template <typename T>
struct NodeMat {
    int N, nf, mom_nnz;
    int ncol;            // nodes of the column space when it differs from the rows' (0: N)
    const int4* info;    // [N] {mom_ptr, cont_ptr, deg, flags}
    const int* nbr_ptr;  // [N+1]
    const int* nbr_idx;  // [graph_nnz]
    const T* vals;       // [nnz]
};

// Node-major sliced ELL copy of a case's FP64 system (rebuilt from the values after every
// assembly) for the hot SpMV-type kernels: nodes in slices of 64, one wavefront lane per
// node, each slice padded to its largest node degree W_c.  A node's d+1 rows are computed
// by its own lane: no cross-lane sums, and every value load is 64 consecutive doubles.
// Rows holding only their diagonal keep it at (rank of the node itself, t = s) and zeros
// elsewhere, so every row is computed the same way.
constexpr int kEllSlice = 64;  // nodes per slice: one wavefront

template <typename VT>
struct EllMatT {
    int N, nf, nsl;       // row nodes, fields, slices
    int ncol;             // column nodes when they differ from the rows' (0: N)
    const int* soff;      // [nsl + 1] sum of the widths of the earlier slices
    const int* idx;       // [soff[nsl] * 64] column node of lane r at rank k: idx[(soff[c] + k) * 64 + r]
                          // (padding ranks: the node itself or node 0, with zero values)
    const VT* vals;       // [soff[nsl] * nf * nf * 64] value (s, t) of lane r at rank k:
                          // vals[((soff[c] + k) * nf * nf + s * nf + t) * 64 + r]; null: no copy
};
using EllMat = EllMatT<double>;

// Scalars of one BiCGSTAB solve.  The per-iteration scalars live in two slots used
// alternately (the host passes the slot parity of each iteration): a kernel only writes
// the current slot and only reads the previous one or fields written by earlier kernels,
// so no workgroup ever reads a value another workgroup of the same launch is writing.
struct KSlot {
    double rho, alpha, omega, pad;
};
struct KScal {
    KSlot slot[2];
    double rr;    // last ||r||^2
    double bb;    // ||b||^2
    double tol2;  // absolute squared tolerance on ||r||^2
    double rtol;
    double atol2; // > 0: ||r||^2 must also fall below it (absolute cap), 0: relative only
    int done;     // 0 running, 1 converged, 2 breakdown or non-finite, 3 iteration limit
    int iters;
    int maxit;
    int req_init;  // set by the host: the next init launch (re)starts this case
    int true_rr;   // 1 when rr is a true residual b - A x (set by init, cleared by each iteration)
    int pad;
};

// Slots of the per-workgroup partial sums (kPartStride doubles per workgroup).
enum : int { P_RHR = 0, P_RR = 1, P_RV = 2, P_SS = 3, P_TS = 4, P_TT = 5, P_BB = 6, kPartStride = 8 };

// BiCGSTAB workspace of one case; T is the inner arithmetic type (double or float).
// Rows are n = nf * nodes (field-blocked); workgroups are assigned XCD-local node ranges.
template <typename T>
struct KrylovView {
    int n, x_zero;
    int nodes, nf;
    int nblk;       // workgroups of this case in the Krylov launches (partials producers)
    int amg;        // 1: ph and sh come from the multigrid V-cycle (K_p / K_s leave them)
    NodeMat<T> A;
    EllMatT<T> E;   // the ELL copy of A in type T used by K_v / K_t (vals null: node-block A)
    const T* binv;  // [nf*nf][nodes] inverse of each node's diagonal (nf x nf) block, SoA
    const T* b;
    T* x;
    T* r;
    T* rh;
    T* p;
    T* v;
    T* s;
    T* t;
    T* ph;
    T* sh;
    double* partials;    // [kMaxCaseBlocks * kPartStride]
    unsigned* counters;  // [8]
    KScal* sc;
    // FP64 multigrid: K_p / K_s also write the V-cycle's level-0 pre-smoothing w = om B^-1 p
    // (or s) from the vector they hold (null: the V-cycle's own pre-smoothing launch does it)
    double* mg_w0;
    double mg_om;
};

// One level of a case's aggregation multigrid (amg.hpp) as seen by the setup and V-cycle
// kernels; T is the hierarchy's arithmetic type (FP64, or FP32 under FP64 BiCGSTAB: the
// preconditioner only shapes the search directions, the solve converges on the FP64 system).
constexpr int kAmgMaxDense = 88;  // coarsest-level unknowns inverted densely in LDS
constexpr int kRestrictLanes = 8;  // lanes per aggregate of the restriction
// Lanes per node in the PSPG assembly gathers (pair blocks, diagonal block and b).  Measured on
// the 262k-node 3D dam break / a 100k-node 2D column, off-diagonal + diagonal kernels: 16 lanes
// 767 + 370 / 41 + 44 us, 8 lanes 747 + 298 / 34 + 29 us, 4 lanes 888 + 298 / 49 + 27 us.
constexpr int kAsmLanes = 8;

template <typename T>
struct AmgView {
    NodeMat<T> A;          // this level's matrix (level 0: the system or its FP32 copy)
    EllMatT<T> E0;         // level 0: the system's ELL copy in type T (vals null: none)
    EllMatT<float> E0f;    // level 0 of an FP64 hierarchy: FP32 values read by the residual (vals null: none)
    const T* binv;         // block-Jacobi inverse, SoA [(nf*nf)][N]
    int level;
    int lanes;             // SpMV lanes per node on this level
    int mask;              // level 0: transfers and Galerkin sums use full rows only
    const uint8_t* flags;  // level 0: the nodes' flags (the masks read them; 1 byte per node)
    int nc;                // coarsest level: unknowns of the dense inverse
    double omega;          // block-Jacobi damping of the smoother
    double alpha;          // scaling of the coarse-grid correction (over-correction of the
                           // piecewise-constant prolongation)
    const int* done_flag;  // the case's BiCGSTAB done flag (finished solves skip the cycle)
    const double* r0[2];   // level 0 inputs (p, s of the FP64 BiCGSTAB)
    double* out0[2];       // level 0 outputs (ph, sh)
    T* r;                  // level >= 1 input
    T* w;                  // smoothed iterate
    T* t;                  // residual
    T* z;                  // level >= 1 output
    // transfer to the next coarser level
    int next_N;
    int pad0;
    const int* agg_ptr;    // members (nodes of this level) of each next-level aggregate
    const int* agg_nodes;
    const int* fine_agg;   // next-level aggregate of each node of this level (-1: none)
    T* next_r;
    const T* next_z;
    // A P in the sliced ELL layout: this level's rows, the next coarser level's columns; the
    // entry block (i, J) sums the blocks (i, j) of i's neighbours j in aggregate J (level 0:
    // over full rows (j, t) only; zero in rows (i, s) that hold only their diagonal)
    EllMatT<T> AP;
    int ap_lanes;          // lanes per node of the set-up kernel
    int pad1;
    const int* ap_ptr;     // [N + 1] AP entries of each node (ranks 0 .. degree-1 of its ELL row)
    const int* ap_gptr;    // [AP entries + 1] gather list of each AP entry
    const int* ap_gsrc;    // rank k in i's neighbour list of each gathered block
    T* ap_out;             // AP values (written at set-up, padding ranks included)
    EllMatT<float> APf;    // level 0 of an FP64 hierarchy: AP held in FP32 instead (vals null: none)
    float* ap_out32;       // its values (written at set-up in place of ap_out)
    // setup of this level (>= 1) from the finer level: entry block (I, J) is the sum of the
    // finer level's A P blocks (i, J) over the members i of aggregate I
    EllMatT<T> FAP;        // the finer level's A P
    EllMatT<float> FAPf;   // ... or its FP32 copy (level 1 of an FP64 hierarchy whose level 0
                           // holds A P in FP32; vals null: none)
    const int* gal_ptr;    // [entries + 1] gather list of each entry of this level
    const int2* gal_src;   // {finer node i, rank m of J in i's A P row}
    const int* self_rank;
    T* vals_out;
    T* binv_out;
    double* dense;  // coarsest: [nc * nc] inverse, row major over unknowns I + s N (FP64 in
                    // either type: the coarse operators are the worst conditioned)
};

// Mixed-precision iterative refinement (FP64 outer residual, FP32 inner solve).
struct RefineView {
    int n, pad;
    NodeMat<double> A;   // FP64 system
    const double* b;
    double* x;           // FP64 solution
    double* r;           // FP64 residual
    float* r32;          // inner right-hand side
    const float* d32;    // inner correction
    double* partials;
    unsigned* counters;
    double* out;         // [4] {||r||^2, ||b||^2, ...}
};

// Lanes per node of the weakly compressible node-gather passes (2D, 3D): a P1 node has
// ~6 incident triangles in 2D and ~24 tetrahedra in 3D.  On the 2.06M-node 3D tank the
// momentum / continuity gathers take 991 / 284 us at 16 lanes, 830 / 255 at 8 and 919 / 312
// at 4 (whole step 566, 626 and 572 steps/s).
constexpr int kWcLanes[2] = {8, 8};

// Weakly compressible explicit step (WCompNewton/Solver.cpp:236-276).
template <typename S>
struct WcView {
    int dim, N, E, pad;
    const int* conn;          // [ne][E]
    const uint8_t* flags;     // [N] with NODE_FREE derived
    const int* elem_ptr;      // [N+1] node-element incidence (Pattern::elem_ptr / elem_list /
    const uint32_t* elem_list;  //       elem_nodes)
    const int4* elem_nodes;
    double* pos;              // [dim][N]
    S* v;                     // [dim][N]
    S* a;                     // [dim][N]
    S* p;                     // [N]
    S* rho;                   // [N]
    const double* dir_vals;   // [dim][N]
    double* m_cont;           // [N] lumped continuity mass of each node
    double* f_cont;           // [N] continuity rhs
    double* m_mom;            // [N] lumped momentum mass (same for every component)
    double* f_mom;            // [dim][N] momentum force
    double* hmin;             // [N] smallest he of each node's elements (CFL after the step)
    double* rec;              // [N][8] node records {x, y, z, p, vx, vy, vz, rho} (FP64; 2D: z = vz = 0),
                              // written by the node passes, read by the element gathers
    double* erec;             // [E][4] continuity element records {|e|, div v, sum P, 0} (FP64)
    unsigned long long* dt2;  // CFL min accumulator (bits of a positive double)
    double mu, K0, K0p, rho_star;
    double g[3];
    double dt;   // time step of the current launch
    double sc2;  // securityCoeff^2 of the CFL reduction
    int meduri, pad1;
    FstView fst;  // surface tension (fst.F == 0: none)
};

// Picard loop helpers (MomContEquationPSPG.inl:293-372).
struct PicardView {
    int dim, N, n, res_kind;  // res_kind: 0 U, 1 U_P, 2 Ax_f
    double dt;
    const uint8_t* flags;
    double* pos;
    double* pos_save;
    double* q_prev;  // state at the start of the step (committed when the loop converges)
    double* x;       // Picard iterate q
    double* x_prev;  // previous iterate (U / U_P residuals)
    NodeMat<double> A;
    const double* b;
    double* partials;
    unsigned* counters;
    double* out;  // [4]
};

// Bookkeeping copies of the Picard step, one bit per copy (BlockTask::pad of the copy table).
enum : int {
    COPY_PREP = 1,     // pos_save = pos, x = q_prev, x_prev = q_prev   (prepare callback)
    COPY_SAVEIT = 2,   // x_prev = x                                   (next Picard iteration)
    COPY_COMMIT = 4,   // q_prev = x                                   (converged step)
    COPY_RESTORE = 8,  // pos = pos_save                               (Mesh::restoreNodesList)
    COPY_ZERO_X = 16,  // x = 0                                        (restart from zero)
    COPY_RESET_X = 32, // x = q_prev                                   (node states restored)
};

// Scalars of one case in an engine slot, read back by the host once per tick.
struct SlotScal {
    KScal k[2];              // FP64 solve, FP32 inner solve
    double out[4];           // residual sums (Picard residual, refinement residual)
    unsigned long long dt2;  // CFL minimum (bits of a positive double)
    double pad[3];
};

constexpr int kMaxCaseBlocks = 2048;

}  // namespace pfem
