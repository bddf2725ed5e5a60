// Aggregation multigrid hierarchy of a PSPG case (the structure; values are rebuilt on the
// device at every assembly).
//
// Level 0 is the case's own system in its node-block layout.  Level l+1 aggregates the
// nodes of level l: each aggregate is a root node with its unaggregated neighbours, and
// the nodes left over join a neighbouring aggregate.  The transfer is piecewise constant
// per field: coarse DOF (I, s) is the sum (restriction) / the copy (prolongation) of the
// DOFs (i, s) of the nodes i of aggregate I.  At the level-0 -> 1 transfer only DOFs whose
// row is a full assembled row take part (identity rows of boundary and free nodes are left
// to the smoother).  A_l P (finer rows, coarse columns) holds in its block (i, J) the sum of
// the fine blocks (i, j) over i's neighbours j in J, gathered through a precomputed list (no
// atomics, fixed order).  The coarse operators are the Galerkin products A_{l+1} = P^T (A_l P):
// block (I, J) is the sum of the A P blocks (i, J) over the members i of I, gathered the same
// way (a node's neighbours fall in a few aggregates, so this reads about a third of the fine
// blocks a direct sum over (i, j) would).  The post-smoothing uses A P too, so that the
// prolongated correction's residual A (w + P z) = A w + (A P) z costs a product with A P
// instead of a second product with A.
#pragma once

#include <vector>

#include "hostutil.hpp"
#include "pattern.hpp"

namespace pfem {

constexpr int kAmgMaxLevels = 8;

// One coarse level (index >= 1) on the device.
struct AmgLevel {
    int N = 0;          // nodes (aggregates of the finer level)
    int graph_nnz = 0;  // node-graph entries, self included
    int max_deg = 0;
    int lanes = 16;     // SpMV lanes per node
    long long nnz = 0;  // (nf)^2 * graph_nnz values (every row full)
    int mom_nnz = 0;    // nf * graph_nnz
    DevBuf nbr_ptr;     // int  [N+1]
    DevBuf nbr_idx;     // int  [graph_nnz]
    DevBuf node_info;   // int4 [N] {mom_ptr, cont_ptr, deg, flags = 0}
    DevBuf self_rank;   // int  [N] rank of the node in its own neighbour list
    DevBuf agg_ptr;     // int  [N+1]   members of each aggregate (finer-level nodes)
    DevBuf agg_nodes;   // int  [members]
    DevBuf fine_agg;    // int  [N_finer] aggregate of each finer node (-1: not aggregated)
    DevBuf gal_ptr;     // int  [graph_nnz+1] Galerkin gather list of each coarse entry
    DevBuf gal_src;     // int2 [ap_nnz] {finer node i, rank m of the entry's column aggregate
                        //      in i's A P row}
    // A_finer P: rows of the finer level, columns of this level's aggregates
    int ap_nnz = 0;     // entries (finer node, aggregate)
    int ap_lanes = 4;
    DevBuf ap_ptr;      // int  [N_finer+1] entries of each finer node (its distinct adjacent
                        //      aggregates, sorted; ELL ranks 0 .. degree-1)
    DevBuf ap_gptr;     // int  [ap_nnz+1] finer neighbours of each entry
    DevBuf ap_gsrc;     // int  [finer graph entries] rank k in the finer node's list
    // sliced ELL layout of A_finer P (EllMatT): slices of 64 finer nodes
    int ap_nsl = 0;
    long long ap_ranks = 0;
    DevBuf ap_soff;     // int [ap_nsl + 1]
    DevBuf ap_eidx;     // int [ap_ranks * 64] aggregate of each rank (padding: 0, zero values)
};

struct AmgHierarchy {
    int levels = 0;  // including level 0; 0 = no hierarchy
    std::vector<AmgLevel> lev;  // lev[0] unused (level 0 is the case's pattern)
    int coarse_n = 0;           // unknowns of the coarsest level (dense inverse)
};

// Builds the hierarchy of `pat` (already built on the device) on the host and uploads it.
// Coarsening stops at <= coarse_target nodes, when an aggregation pass removes less than
// a fifth of the nodes, or at kAmgMaxLevels.
void build_amg(const Pattern& pat, int dim, int coarse_target, hipStream_t stream, AmgHierarchy& H);

}  // namespace pfem
