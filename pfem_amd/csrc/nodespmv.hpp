// Node-block SpMV of the PSPG matrix (the body of every SpMV-type kernel).
//
// The L lanes of a group compute the nf rows of one node i.  Lane k takes neighbour
// nbr_k of i: it loads the nf*nf values of rows (i, s) at columns (nbr_k, t) — a full
// row holds column (nbr_k, t) at offset t*deg + k, so for a fixed (s, t) the lanes read
// consecutive values — gathers x at (nbr_k, t) once for all nf rows, and accumulates.
// The value loads do not depend on the neighbour index and are issued before the
// index -> x gather chain, so each lane keeps nf*nf + nf loads in flight.  Rows holding
// only their diagonal (momentum rows of boundary nodes, every row of a free node) are
// computed by lane 0.  The group's nf sums are DPP butterflies (group_sum), which keep
// the reduction in the VALU instead of ds_bpermute round trips through the LDS crossbar
// (the same speed on the C2 matrix, no LDS traffic).  Compared with a row-wise CSR SpMV
// this reads the column structure from the node graph (deg indices per node instead of
// nf*nf*deg) and gathers each x entry once per node instead of once per row.  (Measured on
// gfx950, C2 matrix of 516 MB: this kernel 115 us per launch; one node per 64-lane
// wavefront with the fields spread over lanes ~1.6x slower; a wave-sliced layout with fully
// coalesced value loads 118 us; a node-major sliced ELL layout with one lane per node and no
// cross-lane sum 98 us; a plain linear read of the same bytes 75 us.)
#pragma once

#include "common.hpp"
#include "views.hpp"

namespace pfem {

constexpr int kMaxNf = 4;

// y[s] = (A x) of row (i, s), s < nf; valid in every lane of the group.  The values may
// be stored in a narrower type VT than the vectors (FP32 values, FP64 vectors and sums).
template <int L, typename VT, typename T>
__device__ __forceinline__ void node_spmv(const NodeMat<VT>& M, const T* __restrict__ x, long long i, int lane,
                                          T (&y)[kMaxNf]) {
    const int4 inf = M.info[i];
    const int nb = M.nbr_ptr[i];
    const int deg = inf.z, nf = M.nf, D = nf - 1;
    const long long N = M.N, Nx = M.ncol ? M.ncol : M.N;
    const uint8_t f = (uint8_t)inf.w;
    const bool mom_full = (f & (NODE_BOUND | NODE_FREE)) == 0, cont_full = (f & NODE_FREE) == 0;
    long long rs[kMaxNf];
    bool full[kMaxNf];
#pragma unroll
    for (int s = 0; s < kMaxNf; ++s) {
        rs[s] = (s < D) ? (long long)s * M.mom_nnz + inf.x : (long long)D * M.mom_nnz + inf.y;
        full[s] = (s < D) ? mom_full : (s == D ? cont_full : false);
        y[s] = T(0);
    }
    for (int k = lane; k < deg; k += L) {
        T a[kMaxNf][kMaxNf];
#pragma unroll
        for (int s = 0; s < kMaxNf; ++s)
#pragma unroll
            for (int t = 0; t < kMaxNf; ++t)
                a[s][t] = (full[s] && t < nf) ? (T)M.vals[rs[s] + k + (long long)t * deg] : T(0);
        const long long j = M.nbr_idx[nb + k];
        T xv[kMaxNf];
#pragma unroll
        for (int t = 0; t < kMaxNf; ++t) xv[t] = (t < nf) ? x[j + t * Nx] : T(0);
#pragma unroll
        for (int s = 0; s < kMaxNf; ++s) {
            T acc = y[s];
#pragma unroll
            for (int t = 0; t < kMaxNf; ++t) acc += a[s][t] * xv[t];
            y[s] = acc;
        }
    }
    if (lane == 0) {
#pragma unroll
        for (int s = 0; s < kMaxNf; ++s)
            if (s < nf && !full[s]) y[s] = (T)M.vals[rs[s]] * x[i + s * N];
    }
#pragma unroll
    for (int s = 0; s < kMaxNf; ++s)
        if (s < nf) y[s] = group_sum<L>(y[s]);
}

// y[s] for a runtime s without dynamic register indexing.
template <typename T>
__device__ __forceinline__ T pick(const T (&y)[kMaxNf], int s) {
    return s == 0 ? y[0] : s == 1 ? y[1] : s == 2 ? y[2] : y[3];
}

}  // namespace pfem
