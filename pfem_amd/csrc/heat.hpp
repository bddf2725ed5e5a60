// Heat (conduction) step on the GPU: the view the kernels of heat.hip take and their launchers.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

namespace pfem {

struct HeatView {
    int dim = 2, N = 0, E = 0, F = 0, nb = 0;  // nb: workgroups over the nodes (partial-sum slots)
    const int* conn = nullptr;      // [dim+1][E] structure of arrays
    const uint8_t* flags = nullptr; // [N] caller flags, NODE_FREE derived
    uint8_t dbit = 16;              // flag of the Dirichlet temperatures (NODE_DIRICHLET, or NODE_DIRICHLET_T
                                    // in a Boussinesq case, whose NODE_DIRICHLET marks the velocities)
    const double* pos = nullptr;    // [dim][N]
    const int* nbr_ptr = nullptr;   // node graph (self included, sorted)
    const int* nbr_idx = nullptr;
    const int* pei_ptr = nullptr;   // pair-element incidence of every node-graph entry
    const unsigned* pei = nullptr;
    const int* row_ptr = nullptr;   // scalar CSR: full rows = the node's graph entries, else the diagonal
    const int* col = nullptr;
    const int* diag_pos = nullptr;
    double* vals = nullptr;
    double* b = nullptr;
    double* erec = nullptr;         // [E][16]
    const double* Tprev = nullptr;  // node temperatures at the step's start
    const double* Tbc = nullptr;    // Dirichlet temperatures ("<Tag>T"(pos, t + dt))
    const int* facets = nullptr;    // [F][dim]
    const int* fout = nullptr;      // [F]
    const uint8_t* fkind = nullptr; // [F] PFEM_HEAT_* bits
    const double* qgp = nullptr;    // [F][3][dim] or null
    double* fterm = nullptr;        // [F][dim][3] qn, convection, radiation
    const int* nf_ptr = nullptr;    // node -> facet entries (f * dim + local), by facet
    const int* nf_list = nullptr;
    double rho = 0, cv = 0, k = 0, h = 0, Tinf = 0, epsRad = 0, dt = 0;
    // CG
    double* x = nullptr;
    double* r = nullptr;
    double* z = nullptr;
    double* p = nullptr;
    double* q = nullptr;
    double* dinv = nullptr;
    double* part = nullptr;  // [5][nb]: r.z (two parities), r.r, p.q, b.b
    double* scal = nullptr;  // [2]: r.r, b.b after the last published pass
};

int heat_blocks(int n);
void launch_heat_assemble(const HeatView& V, hipStream_t s);
void launch_cg_init(const HeatView& V, hipStream_t s);
void launch_cg_iteration(const HeatView& V, int par, hipStream_t s);

}  // namespace pfem
