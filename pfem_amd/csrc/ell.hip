// Node-major sliced ELL copy of the PSPG system (EllMat in views.hpp): the structure is
// built once per mesh from the node graph; the assembly gathers and the boundary conditions
// (pspg.hip) write its values beside the node-block (CSR-order) ones.
//
// Measured on gfx950 with the C2 matrix (516 MB): the node-block SpMV, L lanes per node
// with a cross-lane sum, runs 115 us per launch; this layout, one lane per node, 98 us
// (2D, 900k nodes: 130 -> 103 us), and 90 us with non-temporal value loads.
#include <hipcub/hipcub.hpp>

#include "common.hpp"
#include "pattern.hpp"
#include "views.hpp"

namespace pfem {

namespace {

constexpr int kSlice = kEllSlice;

// width[c] = the largest degree of slice c's nodes, at least 1: a free node has no graph
// entry (degree 0) but keeps its diagonal at rank 0; width[nsl] = 0 (closes the scan).
__global__ void k_ell_width(int N, int nsl, const int* __restrict__ nbr_ptr, int* __restrict__ width) {
    for (int c = blockIdx.x * blockDim.x + threadIdx.x; c <= nsl; c += gridDim.x * blockDim.x) {
        int w = 0;
        if (c < nsl)
            for (int r = 0; r < kSlice; ++r) {
                const int i = c * kSlice + r;
                if (i < N) w = max(w, max(1, nbr_ptr[i + 1] - nbr_ptr[i]));
            }
        width[c] = w;
    }
}

// One thread per (slice, lane): the column node of every rank (padding: the node itself;
// lanes past the last node: node 0, with zero values).
__global__ void k_ell_idx(int N, int nsl, const int* __restrict__ nbr_ptr, const int* __restrict__ nbr_idx,
                          const int* __restrict__ soff, int* __restrict__ idx) {
    const long long total = (long long)nsl * kSlice;
    for (long long q = blockIdx.x * (long long)blockDim.x + threadIdx.x; q < total; q += (long long)gridDim.x * blockDim.x) {
        const int c = (int)(q / kSlice), r = (int)(q % kSlice);
        const int i = c * kSlice + r;
        const long long b = soff[c];
        const int W = soff[c + 1] - (int)b;
        const int beg = i < N ? nbr_ptr[i] : 0, deg = i < N ? nbr_ptr[i + 1] - beg : 0;
        for (int k = 0; k < W; ++k) idx[(b + k) * kSlice + r] = k < deg ? nbr_idx[beg + k] : (i < N ? i : 0);
    }
}

}  // namespace

void build_ell(Pattern& pat, hipStream_t stream) {
    const int N = pat.N;
    const int nsl = (N + kSlice - 1) / kSlice;
    DevBuf width, tmp;
    width.reserve(sizeof(int) * ((size_t)nsl + 1));
    pat.ell_soff.reserve(sizeof(int) * ((size_t)nsl + 1));
    k_ell_width<<<grid_for(nsl + 1, kBlock, 65536), kBlock, 0, stream>>>(N, nsl, pat.nbr_ptr.as<int>(), width.as<int>());
    size_t bytes = 0;
    PFEM_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, bytes, width.as<int>(), pat.ell_soff.as<int>(), nsl + 1, stream));
    tmp.reserve(bytes);
    PFEM_HIP(hipcub::DeviceScan::ExclusiveSum(tmp.p, bytes, width.as<int>(), pat.ell_soff.as<int>(), nsl + 1, stream));
    int ranks = 0;
    PFEM_HIP(hipMemcpyAsync(&ranks, pat.ell_soff.as<int>() + nsl, sizeof(int), hipMemcpyDeviceToHost, stream));
    PFEM_HIP(hipStreamSynchronize(stream));
    pat.ell_nsl = nsl;
    pat.ell_ranks = ranks;
    pat.ell_idx.reserve(sizeof(int) * (size_t)ranks * kSlice);
    k_ell_idx<<<grid_for((long long)nsl * kSlice, kBlock, 65536), kBlock, 0, stream>>>(
        N, nsl, pat.nbr_ptr.as<int>(), pat.nbr_idx.as<int>(), pat.ell_soff.as<int>(), pat.ell_idx.as<int>());
    PFEM_HIP(hipStreamSynchronize(stream));
}


}  // namespace pfem
