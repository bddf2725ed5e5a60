// Fractional-step (projection) Picard step on the GPU: the views the kernels of fracstep.hip
// take and their launchers.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

namespace pfem {

// The step's three matrices share the node graph: value k of node i's row is graph entry k
// (node pair (i, nbr_idx[k])).  Rows of bound or free nodes are unit rows (MK_dt, M), columns
// of bound or free nodes are unit columns in M, Dirichlet columns of MK_dt are eliminated.
struct FracView {
    int dim = 2, N = 0, E = 0, nb = 0;  // nb: workgroups over the nodes (residual partial slots)
    const int* conn = nullptr;           // [dim+1][E]
    const uint8_t* flags = nullptr;      // [N] caller flags, NODE_FREE derived from connectivity
    double* pos = nullptr;               // [dim][N] current positions
    const double* pos_save = nullptr;    // [dim][N] positions at the step's start
    const int* nbr_ptr = nullptr;        // node graph (sorted, self included; a free node has no entry)
    const int* nbr_idx = nullptr;
    const int* pei_ptr = nullptr;        // pair-element incidence of every node-graph entry
    const unsigned* pei = nullptr;
    double* erec = nullptr;              // [E][16]: dN_i/dx_k at k (d+1) + i, K factor at 14, |e| at 15
    double* mk = nullptr;                // [graph_nnz][dim][dim]  M/dt + K (velocity prediction)
    double* ms = nullptr;                // [graph_nnz]            M (velocity correction, per component)
    double* lv = nullptr;                // [graph_nnz]            L (pressure correction)
    double* dinv_mk = nullptr;           // [dim][N]
    double* dinv_m = nullptr;            // [dim][N] (the same per component)
    double* dinv_l = nullptr;            // [N]
    double* bv = nullptr;                // [dim][N] right-hand side of the two velocity systems
    double* bp = nullptr;                // [N]      right-hand side of the pressure system
    const double* v_prev = nullptr;      // [dim][N] velocity at the step's start
    const double* p_prev = nullptr;      // [N]      pressure at the step's start
    const double* vbc = nullptr;         // [dim][N] Dirichlet velocities
    double* vt = nullptr;                // [dim][N] predicted velocity (v tilde)
    double* pn = nullptr;                // [N]      pressure of the iterate
    double* dv = nullptr;                // [dim][N] velocity correction
    double* v_it = nullptr;              // [dim][N] velocity of the iterate (the node states)
    double* v_itp = nullptr;             // [dim][N] velocity of the previous iterate
    double* p_itp = nullptr;             // [N]      pressure of the previous iterate
    double* res_part = nullptr;          // [4][nb]  residual partial sums
    double rho = 0, mu = 0, dt = 0, gamma_fs = 0, tau0 = 0, mreg = 0;
    double g[3] = {0, 0, 0};
    int problem = 0;                     // 1: Bingham K factor
};

// Diagonal-preconditioned conjugate gradient on one of the three systems.
struct CgView {
    int op = 0;       // 0: scalar graph matrix on N rows; 1: scalar graph matrix applied to each of
                      // dim components (dim N rows); 2: dim x dim block graph matrix (dim N rows)
    int dim = 2, N = 0, n = 0, nb = 0;
    const int* nbr_ptr = nullptr;
    const int* nbr_idx = nullptr;
    const double* vals = nullptr;
    const double* b = nullptr;
    const double* dinv = nullptr;
    double* x = nullptr;
    double* r = nullptr;
    double* z = nullptr;
    double* p = nullptr;
    double* q = nullptr;
    double* part = nullptr;  // [5][nb]: r.z (two parities), r.r, p.q, b.b
    double* scal = nullptr;  // [2]: r.r, b.b after the last published pass
};

int frac_blocks(int n);
// element records, the three matrices, the prediction right-hand side and its boundary conditions
void launch_frac_assemble(const FracView& V, hipStream_t s);
// right-hand side of the pressure correction (needs V.vt)
void launch_frac_pcorr(const FracView& V, hipStream_t s);
// right-hand side of the velocity correction (needs V.pn)
void launch_frac_vstep(const FracView& V, hipStream_t s);
// iterate = vt + dv, positions, residual partials; previous iterate <- iterate
void launch_frac_update(const FracView& V, hipStream_t s);
void launch_cg2_init(const CgView& C, hipStream_t s);
void launch_cg2_iteration(const CgView& C, int par, hipStream_t s);

}  // namespace pfem
