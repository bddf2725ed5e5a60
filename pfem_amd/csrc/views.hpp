// Plain-old-data views of one case's device arrays, passed to kernels through device
// arrays indexed by BlockTask::case_id.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

namespace pfem {

// PSPG assembly + boundary conditions (MomContEquationPSPG.inl:7-235).
struct PspgView {
    int dim, N, E, mom_nnz;
    int n_dir, pad0;
    const int* conn;          // [ne][E]
    const double* pos;        // [dim][N]
    const double* q_prev;     // [(dim+1)N]  states at the start of the time step (be, free-node rhs)
    const double* q_iter;     // [(dim+1)N]  current Picard iterate (node states used by tau)
    const int4* node_info;    // [N] {mom_ptr, cont_ptr, deg, flags}
    const uint4* ranks;       // [E] packed 8-bit slot ranks
    const int4* dir;          // [n_dir] {position, row, column, 0}
    const double* dir_vals;   // [dim][N] Dirichlet velocities (read at Dirichlet nodes)
    double* vals;             // [nnz]
    double* b;                // [(dim+1)N]
    float* vals32;            // [nnz] FP32 copy for the mixed-precision inner solve (may be null)
    const int* diag_pos;      // [n]
    double* dinv;             // [n] Jacobi inverse diagonal (FP64)
    float* dinv32;            // [n] (may be null)
    long long nnz;
    double rho, mu, dt;
    double g[3];
};

// Scalars of one BiCGSTAB solve, finalised by the last workgroup of each reduction.
struct KScal {
    double rho, alpha, omega, beta;
    double rr, bb, ss, rv;
    double tol2;  // absolute squared tolerance on ||r||^2
    double rtol;
    int done;     // 0 running, 1 converged, 2 breakdown or non-finite, 3 iteration limit
    int iters;
    int maxit;
    int s_conv;
    int req_init;  // set by the host: the next init launch (re)starts this case
    int pad;
};

// BiCGSTAB workspace of one case; T is the inner arithmetic type (double or float).
template <typename T>
struct KrylovView {
    int n, x_zero;
    const int* row_ptr;
    const int* col_idx;
    const T* vals;
    const T* dinv;
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
    double* partials;    // [kMaxCaseBlocks * 4]
    unsigned* counters;  // [8]
    KScal* sc;
};

// Mixed-precision iterative refinement (FP64 outer residual, FP32 inner solve).
struct RefineView {
    int n, pad;
    const int* row_ptr;
    const int* col_idx;
    const double* vals;  // FP64 system
    const double* b;
    double* x;           // FP64 solution
    double* r;           // FP64 residual
    float* r32;          // inner right-hand side
    const float* d32;    // inner correction
    double* partials;
    unsigned* counters;
    double* out;         // [4] {||r||^2, ||b||^2, ...}
};

// Weakly compressible explicit step (WCompNewton/Solver.cpp:236-276).
template <typename S>
struct WcView {
    int dim, N, E, pad;
    const int* conn;          // [ne][E]
    const uint8_t* flags;     // [N] with NODE_FREE derived
    double* pos;              // [dim][N]
    S* v;                     // [dim][N]
    S* a;                     // [dim][N]
    S* p;                     // [N]
    S* rho;                   // [N]
    const double* dir_vals;   // [dim][N]
    double* m_cont;           // [N] lumped continuity mass accumulator
    double* f_cont;           // [N] continuity rhs accumulator
    double* m_mom;            // [N] lumped momentum mass accumulator (same for every component)
    double* f_mom;            // [dim][N] momentum force accumulator
    unsigned long long* dt2;  // CFL min accumulator (bits of a positive double)
    double mu, K0, K0p, rho_star;
    double g[3];
    double dt;   // time step of the current launch
    double sc2;  // securityCoeff^2 of the CFL reduction
    int meduri, pad1;
};

// Picard loop helpers (MomContEquationPSPG.inl:293-372).
struct PicardView {
    int dim, N, n, res_kind;  // res_kind: 0 U, 1 U_P, 2 Ax_f
    double dt;
    const uint8_t* flags;
    double* pos;
    const double* pos_save;
    const double* x;       // new iterate q
    const double* x_prev;  // previous iterate (U / U_P residuals)
    const int* row_ptr;
    const int* col_idx;
    const double* vals;
    const double* b;
    double* partials;
    unsigned* counters;
    double* out;  // [4]
};

constexpr int kMaxCaseBlocks = 2048;

}  // namespace pfem
