// Node-major sliced ELL copy of the PSPG system (EllMat in views.hpp): the structure is
// built once per mesh from the node graph, the values are refilled from the node-block
// (CSR-order) values after every assembly.
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

__device__ __forceinline__ int rank_of(const int* __restrict__ nbr_idx, int beg, int deg, int target) {
    int lo = 0, hi = deg;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (nbr_idx[beg + mid] < target) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

// Value (s, t, rank k) of node i, with its row start rs and degree deg, from `row` (the
// node-block values of row (i, s), in LDS or global memory).
__device__ __forceinline__ double ell_value(const double* row, bool full, int deg, int self, int s, int t, int k) {
    if (full) return k < deg ? row[(long long)t * deg + k] : 0.0;
    return (k == self && t == s) ? row[0] : 0.0;
}

constexpr int kFillLds = 8192;  // doubles of one field block of a slice staged in LDS (64 KB)

// One workgroup per slice.  For each field s the slice's rows (i, s) are consecutive in
// the node-block values: they are staged in LDS with coalesced loads, then each wavefront
// writes whole 64-value rank vectors of the copy (lane r = node 64 c + r).  A slice whose
// field block exceeds the LDS stage reads its rows from global memory instead.
__global__ __launch_bounds__(kBlock) void k_ell_fill(const PspgView* __restrict__ views,
                                                     const BlockTask* __restrict__ tasks) {
    __shared__ double stage[kFillLds];
    const BlockTask task = tasks[blockIdx.x];
    if (task.case_id < 0) return;
    const PspgView& V = views[task.case_id];
    if (!V.vals_ell) return;
    const int N = V.N, D = V.dim, nf = D + 1;
    const int lane = threadIdx.x & (kSlice - 1), wave = threadIdx.x >> 6;
    for (long long c = task.block; c < V.ell_nsl; c += task.nblocks) {
        const long long i = c * kSlice + lane;
        const long long b = V.ell_soff[c];
        const int W = V.ell_soff[c + 1] - (int)b;
        const int i0 = (int)(c * kSlice), i1 = min(N, i0 + kSlice) - 1;  // first and last node
        int4 inf = make_int4(0, 0, 0, 0);
        int self = 0;
        if (i < N) {
            inf = V.node_info[i];
            self = rank_of(V.nbr_idx, V.nbr_ptr[i], inf.z, (int)i);
        }
        const uint8_t f = (uint8_t)inf.w;
        const bool mom_full = (f & (NODE_BOUND | NODE_FREE)) == 0, cont_full = (f & NODE_FREE) == 0;
        const int4 inf0 = V.node_info[i0], inf1 = V.node_info[i1];
        for (int s = 0; s < nf; ++s) {
            const bool full = (s < D) ? mom_full : cont_full;
            auto rstart = [&](const int4& q) {
                return (s < D) ? (long long)s * V.mom_nnz + q.x : (long long)D * V.mom_nnz + q.y;
            };
            const uint8_t f1 = (uint8_t)inf1.w;
            const bool full1 = (s < D) ? (f1 & (NODE_BOUND | NODE_FREE)) == 0 : (f1 & NODE_FREE) == 0;
            const long long beg = rstart(inf0), end = rstart(inf1) + (full1 ? (long long)nf * inf1.z : 1);
            const bool staged = end - beg <= kFillLds;
            __syncthreads();  // the previous field's stage is no longer read
            if (staged)
                for (long long q = beg + threadIdx.x; q < end; q += kBlock) stage[q - beg] = V.vals[q];
            __syncthreads();
            const double* row = staged ? stage + (rstart(inf) - beg) : V.vals + rstart(inf);
            for (int p = wave; p < W * nf; p += kWaves) {
                const int k = p / nf, t = p - k * nf;
                const double v = (i < N) ? ell_value(row, full, inf.z, self, s, t, k) : 0.0;
                const long long q = ((b + k) * nf * nf + s * nf + t) * kSlice + lane;
                V.vals_ell[q] = v;
                if (V.vals32_ell) V.vals32_ell[q] = (float)v;
            }
        }
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

void launch_ell_fill(const PspgView* views, const BlockTask* tasks, int ntasks, hipStream_t s) {
    k_ell_fill<<<ntasks, kBlock, 0, s>>>(views, tasks);
}

}  // namespace pfem
