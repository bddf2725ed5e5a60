// SpMV over the sliced ELL copy of a PSPG system (EllMat): lane r of a wavefront computes
// the NF rows of node 64 c + r of slice c.  Each rank k is one column node j (one index
// load, NF gathers of x) and NF*NF value loads, each 64 consecutive doubles per wavefront.
// The values are loaded non-temporally: they are read once per launch, and without the hint
// they evict the x entries and the results from the caches (measured on gfx950, 516 MB
// matrix: 98.7 -> 90.0 us per launch; 2D 900k nodes: 104 -> 94 us).
#pragma once

#include "common.hpp"
#include "nodespmv.hpp"
#include "views.hpp"

namespace pfem {

// y[s] = (A x) of row (64 c + lane, s) for s < NF (y[s] = 0 above), with values of type VT
// and vectors and sums of type XT; lanes past the last node compute values that the
// callers drop.
template <int NF, typename VT, typename XT>
__device__ __forceinline__ void ell_rows(const EllMatT<VT>& E, const XT* __restrict__ x, long long c, int lane,
                                         XT (&y)[kMaxNf]) {
    const long long b = E.soff[c];
    const int W = E.soff[c + 1] - (int)b;
    const VT* __restrict__ vb = E.vals + b * (NF * NF * kEllSlice) + lane;
    const int* __restrict__ ib = E.idx + b * kEllSlice + lane;
    const long long Nx = E.ncol ? E.ncol : E.N;
    XT acc[NF];
#pragma unroll
    for (int s = 0; s < NF; ++s) acc[s] = XT(0);
#pragma unroll 2
    for (int k = 0; k < W; ++k) {
        const long long j = ib[(long long)k * kEllSlice];
        XT a[NF][NF];
#pragma unroll
        for (int s = 0; s < NF; ++s)
#pragma unroll
            for (int t = 0; t < NF; ++t)
                a[s][t] = (XT)ld_stream(vb + ((long long)k * NF * NF + s * NF + t) * kEllSlice);
        XT xv[NF];
#pragma unroll
        for (int t = 0; t < NF; ++t) xv[t] = x[j + t * Nx];
#pragma unroll
        for (int s = 0; s < NF; ++s)
#pragma unroll
            for (int t = 0; t < NF; ++t) acc[s] += a[s][t] * xv[t];
    }
#pragma unroll
    for (int s = 0; s < kMaxNf; ++s) y[s] = s < NF ? acc[s] : XT(0);
}

}  // namespace pfem
