// K0: CSR pattern of the PSPG system, built on the GPU from element connectivity.
#pragma once

#include <cstdint>
#include "hostutil.hpp"

namespace pfem {

// DOF layout follows the reference's field-blocked q vectors (StatesFromToQ.hpp:21-34):
// row/column (node i, field s) is i + s*N, s = 0..dim-1 velocity, s = dim pressure.
//
// Row (i, s) is "full" when the reference emits element triplets into it
// (MomContEquationPSPG.inl:68,81): momentum rows of nodes that are neither boundary
// nor free, continuity rows of non-free nodes.  A full row holds, for t = 0..dim and
// every node j of i's sorted node-graph neighbourhood (i itself included), column
// j + t*N; so its length is (dim+1)*deg(i) and column (j, t) sits at offset
// t*deg(i) + rank_i(j).  Every other row holds only its diagonal (the unit triplets
// pushed at MomContEquationPSPG.inl:108-128).  Row starts are
//   s <  dim: s*mom_nnz + mom_ptr[i],   s == dim: dim*mom_nnz + cont_ptr[i],
// where mom_ptr / cont_ptr are prefix sums of the momentum / continuity row lengths.
struct Pattern {
    int dim = 0, N = 0, E = 0, n = 0;
    long long nnz = 0;
    int mom_nnz = 0, cont_nnz = 0;
    int graph_nnz = 0;  // node-graph entries (node pairs sharing an element, incl. self)
    int max_deg = 0;
    DevBuf conn;       // int  [ne][E]  element connectivity, structure of arrays
    DevBuf flags;      // u8   [N]      node flags with NODE_FREE derived from connectivity
    DevBuf nbr_ptr;    // int  [N+1]
    DevBuf nbr_idx;    // int  [graph_nnz]
    DevBuf node_info;  // int4 [N]      {mom_ptr, cont_ptr, deg, flags}
    DevBuf row_ptr;    // int  [n+1]
    DevBuf diag_pos;   // int  [n]      position of the diagonal entry of each row
    // pair-element incidence: for node-graph entry m = (i, j), the elements containing
    // both nodes, sorted by element index, encoded e << 4 | a << 2 | b with conn[e][a] = i
    // and conn[e][b] = j
    DevBuf pei_ptr;    // int  [graph_nnz + 1]
    DevBuf pei;        // u32  [E * ne * ne]
    // node-element incidence (weakly compressible cases): for node i, entries
    // elem_ptr[i] .. elem_ptr[i+1] hold e << 2 | a with conn[e][a] = i, sorted by element
    DevBuf elem_ptr;   // int  [N + 1]
    DevBuf elem_list;  // u32  [E * ne]
    DevBuf elem_nodes; // int4 [E * ne] the element's nodes (local order) of each entry, so that
                       //      a gather over a node's elements needs no connectivity lookup
    // node-major sliced ELL structure (EllMat in views.hpp; PSPG cases)
    int ell_nsl = 0;          // slices of 64 nodes
    long long ell_ranks = 0;  // sum of the slice widths
    DevBuf ell_soff;          // int [ell_nsl + 1]
    DevBuf ell_idx;           // int [ell_ranks * 64]
};

// Builds every array of `pat` on `stream` (synchronises the stream before returning,
// since array sizes are data dependent).  conn_aos: device, E x (dim+1) row-major;
// flags_in: device, N bytes (NODE_FREE is recomputed).
void build_pattern(int dim, int N, int E, const int* conn_aos, const uint8_t* flags_in, hipStream_t stream,
                   Pattern& pat);

// Builds the sliced ELL structure of `pat` (slice widths and offsets, padded column nodes)
// from its node graph (synchronises the stream).
void build_ell(Pattern& pat, hipStream_t stream);

// Writes the nnz CSR column indices of `pat` into col_idx (device) on `stream` (export
// only: the kernels address values through the node graph).
void build_col_idx(const Pattern& pat, int* col_idx, hipStream_t stream);

// The structure-of-arrays connectivity, the derived node flags and the node-element
// incidence (what the weakly compressible step needs; no matrix).
void build_mesh_basic(int dim, int N, int E, const int* conn_aos, const uint8_t* flags_in, hipStream_t stream,
                      Pattern& pat);

}  // namespace pfem
