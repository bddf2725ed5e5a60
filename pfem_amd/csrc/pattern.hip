// K0: node graph, field-blocked DOF CSR row structure, diagonal positions and the
// pair-element incidence of the gather assembly, all built on the device from the
// element connectivity.
//
// Replaces the pattern side of Eigen's SparseMatrix::setFromTriplets at
// MomContEquationPSPG.inl:136 (the (row, col) set the triplet list of
// MomContEquationPSPG.inl:55-128 produces).  Column indices are not stored: the kernels
// address values through the node graph, and build_col_idx writes them for export.
#include <hipcub/hipcub.hpp>

#include "common.hpp"
#include "pattern.hpp"

namespace pfem {
namespace {

__global__ void k_conn_soa(int E, int ne, int N, const int* __restrict__ aos, int* __restrict__ soa,
                           int* __restrict__ count, int* __restrict__ err) {
    for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < E; e += (long long)gridDim.x * blockDim.x) {
        for (int k = 0; k < ne; ++k) {
            int nd = aos[e * ne + k];
            if (nd < 0 || nd >= N) {
                atomicOr(err, 1);
                nd = 0;
            } else {
                atomicAdd(&count[nd], 1);
            }
            soa[(long long)k * E + e] = nd;
        }
    }
}

__global__ void k_flags(int N, const uint8_t* __restrict__ in, const int* __restrict__ count, uint8_t* __restrict__ out) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < N; i += gridDim.x * blockDim.x) {
        uint8_t f = in[i] & (uint8_t)~NODE_FREE;
        if (count[i] == 0) f |= NODE_FREE;
        out[i] = f;
    }
}

template <int NE>
__global__ void k_pair_keys(int E, int N, const int* __restrict__ soa, unsigned long long* __restrict__ keys) {
    for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < E; e += (long long)gridDim.x * blockDim.x) {
        int nd[NE];
#pragma unroll
        for (int k = 0; k < NE; ++k) nd[k] = soa[(long long)k * E + e];
#pragma unroll
        for (int a = 0; a < NE; ++a)
#pragma unroll
            for (int b = 0; b < NE; ++b)
                keys[e * (NE * NE) + a * NE + b] = (unsigned long long)nd[a] * (unsigned long long)N + nd[b];
    }
}

// nbr_ptr[i] = first unique key >= i*N (lower bound), for i = 0..N.
__global__ void k_nbr_ptr(int N, int M, const unsigned long long* __restrict__ ukeys, int* __restrict__ nbr_ptr) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i <= N; i += gridDim.x * blockDim.x) {
        const unsigned long long target = (unsigned long long)i * (unsigned long long)N;
        int lo = 0, hi = M;
        while (lo < hi) {
            int mid = (lo + hi) >> 1;
            if (ukeys[mid] < target) lo = mid + 1;
            else hi = mid;
        }
        nbr_ptr[i] = lo;
    }
}

__global__ void k_nbr_idx(int M, int N, const unsigned long long* __restrict__ ukeys, int* __restrict__ nbr_idx) {
    for (int m = blockIdx.x * blockDim.x + threadIdx.x; m < M; m += gridDim.x * blockDim.x)
        nbr_idx[m] = (int)(ukeys[m] % (unsigned long long)N);
}

__device__ __forceinline__ bool mom_full(uint8_t f) { return (f & (NODE_BOUND | NODE_FREE)) == 0; }
__device__ __forceinline__ bool cont_full(uint8_t f) { return (f & NODE_FREE) == 0; }

__global__ void k_row_len(int N, int dim, const int* __restrict__ nbr_ptr, const uint8_t* __restrict__ flags,
                          int* __restrict__ len_mom, int* __restrict__ len_cont) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i <= N; i += gridDim.x * blockDim.x) {
        if (i == N) {
            len_mom[N] = 0;
            len_cont[N] = 0;
            continue;
        }
        const int deg = nbr_ptr[i + 1] - nbr_ptr[i];
        const uint8_t f = flags[i];
        len_mom[i] = mom_full(f) ? (dim + 1) * deg : 1;
        len_cont[i] = cont_full(f) ? (dim + 1) * deg : 1;
    }
}

__global__ void k_node_info(int N, const int* __restrict__ mom_ptr, const int* __restrict__ cont_ptr,
                            const int* __restrict__ nbr_ptr, const uint8_t* __restrict__ flags, int4* __restrict__ info,
                            int* __restrict__ max_deg) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < N; i += gridDim.x * blockDim.x) {
        const int deg = nbr_ptr[i + 1] - nbr_ptr[i];
        info[i] = make_int4(mom_ptr[i], cont_ptr[i], deg, (int)flags[i]);
        atomicMax(max_deg, deg);
    }
}

__device__ __forceinline__ long long row_start(const int4 inf, int s, int dim, int mom_nnz) {
    return (s < dim) ? (long long)s * mom_nnz + inf.x : (long long)dim * mom_nnz + inf.y;
}
__device__ __forceinline__ bool row_full(const int4 inf, int s, int dim) {
    return (s < dim) ? mom_full((uint8_t)inf.w) : cont_full((uint8_t)inf.w);
}

__global__ void k_row_ptr(int N, int dim, int mom_nnz, long long nnz, const int4* __restrict__ info,
                          int* __restrict__ row_ptr) {
    const long long n = (long long)(dim + 1) * N;
    for (long long r = blockIdx.x * (long long)blockDim.x + threadIdx.x; r <= n; r += (long long)gridDim.x * blockDim.x) {
        if (r == n) {
            row_ptr[n] = (int)nnz;
            continue;
        }
        const int i = (int)(r % N), s = (int)(r / N);
        row_ptr[r] = (int)row_start(info[i], s, dim, mom_nnz);
    }
}

// One thread per node: the columns of every row of the node (CSR export only; the hot
// kernels read values through the node graph).
__global__ void k_col_node(int N, int dim, int mom_nnz, const int4* __restrict__ info, const int* __restrict__ nbr_ptr,
                           const int* __restrict__ nbr_idx, int* __restrict__ col_idx) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < N; i += gridDim.x * blockDim.x) {
        const int4 inf = info[i];
        const int nb = nbr_ptr[i];
        for (int s = 0; s <= dim; ++s) {
            const long long rs = row_start(inf, s, dim, mom_nnz);
            if (!row_full(inf, s, dim)) {
                col_idx[rs] = i + s * N;
                continue;
            }
            for (int k = 0; k < inf.z; ++k) {
                const int j = nbr_idx[nb + k];
                for (int t = 0; t <= dim; ++t) col_idx[rs + (long long)t * inf.z + k] = j + t * N;
            }
        }
    }
}

__device__ __forceinline__ int find_rank(const int* __restrict__ nbr_idx, int beg, int deg, int target) {
    int lo = 0, hi = deg;
    while (lo < hi) {
        int mid = (lo + hi) >> 1;
        if (nbr_idx[beg + mid] < target) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

__global__ void k_col_diag(int N, int dim, int mom_nnz, const int4* __restrict__ info, const int* __restrict__ nbr_ptr,
                           const int* __restrict__ nbr_idx, int* __restrict__ col_idx, int* __restrict__ diag_pos) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < N; i += gridDim.x * blockDim.x) {
        const int4 inf = info[i];
        const int self_rank = (inf.z > 0) ? find_rank(nbr_idx, nbr_ptr[i], inf.z, i) : 0;
        for (int s = 0; s <= dim; ++s) {
            const long long rs = row_start(inf, s, dim, mom_nnz);
            if (row_full(inf, s, dim)) {
                diag_pos[i + (long long)s * N] = (int)(rs + (long long)s * inf.z + self_rank);
            } else {
                if (col_idx) col_idx[rs] = i + s * N;
                diag_pos[i + (long long)s * N] = (int)rs;
            }
        }
    }
}

template <int NE>
__global__ void k_ranks(int E, const int* __restrict__ soa, const int* __restrict__ nbr_ptr,
                        const int* __restrict__ nbr_idx, uint8_t* __restrict__ ranks, int* __restrict__ err) {
    for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < E; e += (long long)gridDim.x * blockDim.x) {
        int nd[NE];
#pragma unroll
        for (int k = 0; k < NE; ++k) nd[k] = soa[(long long)k * E + e];
        uint8_t r[16];
#pragma unroll
        for (int q = 0; q < 16; ++q) r[q] = 0;
#pragma unroll
        for (int a = 0; a < NE; ++a) {
            const int beg = nbr_ptr[nd[a]], deg = nbr_ptr[nd[a] + 1] - beg;
#pragma unroll
            for (int b = 0; b < NE; ++b) {
                const int rk = find_rank(nbr_idx, beg, deg, nd[b]);
                if (rk > 255 || rk >= deg || nbr_idx[beg + rk] != nd[b]) atomicOr(err, 2);
                r[a * NE + b] = (uint8_t)rk;
            }
        }
        uint4 packed;
        packed.x = r[0] | (r[1] << 8) | (r[2] << 16) | ((unsigned)r[3] << 24);
        packed.y = r[4] | (r[5] << 8) | (r[6] << 16) | ((unsigned)r[7] << 24);
        packed.z = r[8] | (r[9] << 8) | (r[10] << 16) | ((unsigned)r[11] << 24);
        packed.w = r[12] | (r[13] << 8) | (r[14] << 16) | ((unsigned)r[15] << 24);
        reinterpret_cast<uint4*>(ranks)[e] = packed;
    }
}

__device__ __forceinline__ uint32_t rank_at(const uint4 r, int q) {
    const uint32_t w = (q < 4) ? r.x : (q < 8) ? r.y : (q < 12) ? r.z : r.w;
    return (w >> (8 * (q & 3))) & 0xffu;
}

// Pair-element incidence: count, fill (cursor order) and per-list sort by element index.
template <int NE>
__global__ void k_pei_count(int E, const int* __restrict__ soa, const uint4* __restrict__ ranks,
                            const int* __restrict__ nbr_ptr, int* __restrict__ cnt) {
    for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < E; e += (long long)gridDim.x * blockDim.x) {
        const uint4 rk = ranks[e];
#pragma unroll
        for (int a = 0; a < NE; ++a) {
            const int base = nbr_ptr[soa[(long long)a * E + e]];
#pragma unroll
            for (int b = 0; b < NE; ++b) atomicAdd(&cnt[base + rank_at(rk, a * NE + b)], 1);
        }
    }
}

template <int NE>
__global__ void k_pei_fill(int E, const int* __restrict__ soa, const uint4* __restrict__ ranks,
                           const int* __restrict__ nbr_ptr, const int* __restrict__ pei_ptr, int* __restrict__ cursor,
                           uint32_t* __restrict__ pei) {
    for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < E; e += (long long)gridDim.x * blockDim.x) {
        const uint4 rk = ranks[e];
#pragma unroll
        for (int a = 0; a < NE; ++a) {
            const int base = nbr_ptr[soa[(long long)a * E + e]];
#pragma unroll
            for (int b = 0; b < NE; ++b) {
                const int m = base + (int)rank_at(rk, a * NE + b);
                const int pos = pei_ptr[m] + atomicAdd(&cursor[m], 1);
                pei[pos] = ((uint32_t)e << 4) | (uint32_t)(a << 2) | (uint32_t)b;
            }
        }
    }
}

// Node-element incidence entries e << 2 | a at node conn[e][a] (order fixed afterwards by k_pei_sort).
__global__ void k_elem_fill(int E, int ne, const int* __restrict__ soa, const int* __restrict__ ptr,
                            int* __restrict__ cursor, uint32_t* __restrict__ list) {
    for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < E; e += (long long)gridDim.x * blockDim.x)
        for (int a = 0; a < ne; ++a) {
            const int nd = soa[(long long)a * E + e];
            list[ptr[nd] + atomicAdd(&cursor[nd], 1)] = ((uint32_t)e << 2) | (uint32_t)a;
        }
}

// Nodes of the element of each node-element incidence entry (padded with -1 in 2D).
__global__ void k_elem_nodes(long long total, int E, int ne, const int* __restrict__ soa,
                             const uint32_t* __restrict__ list, int4* __restrict__ nodes) {
    for (long long q = blockIdx.x * (long long)blockDim.x + threadIdx.x; q < total; q += (long long)gridDim.x * blockDim.x) {
        const long long e = list[q] >> 2;
        int4 v;
        v.x = soa[e];
        v.y = soa[(long long)E + e];
        v.z = soa[2LL * E + e];
        v.w = ne > 3 ? soa[3LL * E + e] : -1;
        nodes[q] = v;
    }
}

// Sorts each list m (entries ptr[m] .. ptr[m+1]) in increasing order.
__global__ void k_pei_sort(int M, const int* __restrict__ pei_ptr, uint32_t* __restrict__ pei) {
    for (int m = blockIdx.x * blockDim.x + threadIdx.x; m < M; m += gridDim.x * blockDim.x) {
        const int beg = pei_ptr[m], end = pei_ptr[m + 1];
        for (int q = beg + 1; q < end; ++q) {  // insertion sort (lists hold at most a few dozen entries)
            const uint32_t v = pei[q];
            int r = q - 1;
            while (r >= beg && pei[r] > v) {
                pei[r + 1] = pei[r];
                --r;
            }
            pei[r + 1] = v;
        }
    }
}

template <class T>
void exclusive_sum(const T* in, T* out, int n, hipStream_t s, DevBuf& tmp) {
    size_t bytes = 0;
    PFEM_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, bytes, in, out, n, s));
    tmp.reserve(bytes);
    PFEM_HIP(hipcub::DeviceScan::ExclusiveSum(tmp.p, bytes, in, out, n, s));
}

}  // namespace

void build_mesh_basic(int dim, int N, int E, const int* conn_aos, const uint8_t* flags_in, hipStream_t stream,
                      Pattern& pat) {
    PFEM_REQUIRE(dim == 2 || dim == 3, E_INVALID, "dim must be 2 or 3");
    PFEM_REQUIRE(N > 0 && E > 0, E_INVALID, "mesh must have nodes and elements");
    const int ne = dim + 1;
    pat.dim = dim;
    pat.N = N;
    pat.E = E;
    pat.n = (dim + 1) * N;
    PFEM_REQUIRE((long long)E * ne < (1LL << 31) && E < (1 << 30), E_INVALID,
                 "too many elements for int32 incidence indexing");
    DevBuf count, err, cursor, tmp;
    pat.conn.reserve(sizeof(int) * (size_t)E * ne);
    pat.flags.reserve((size_t)N);
    count.reserve(sizeof(int) * ((size_t)N + 1));  // count[N] = 0 closes the scan
    err.reserve(sizeof(int));
    PFEM_HIP(hipMemsetAsync(count.p, 0, sizeof(int) * ((size_t)N + 1), stream));
    PFEM_HIP(hipMemsetAsync(err.p, 0, sizeof(int), stream));
    k_conn_soa<<<grid_for(E, kBlock, 65536), kBlock, 0, stream>>>(E, ne, N, conn_aos, pat.conn.as<int>(),
                                                                  count.as<int>(), err.as<int>());
    k_flags<<<grid_for(N, kBlock, 65536), kBlock, 0, stream>>>(N, flags_in, count.as<int>(), pat.flags.as<uint8_t>());
    int herr = 0;
    PFEM_HIP(hipMemcpyAsync(&herr, err.p, sizeof(int), hipMemcpyDeviceToHost, stream));
    PFEM_HIP(hipStreamSynchronize(stream));
    PFEM_REQUIRE(herr == 0, E_INVALID, "connectivity references a node index outside [0, N)");
    // node-element incidence: counts -> offsets, fill through per-node cursors, then sort each
    // node's list by element so that every sum over it has a fixed order
    pat.elem_ptr.reserve(sizeof(int) * ((size_t)N + 1));
    pat.elem_list.reserve(sizeof(uint32_t) * (size_t)E * ne);
    cursor.reserve(sizeof(int) * (size_t)N);
    PFEM_HIP(hipMemsetAsync(cursor.p, 0, sizeof(int) * (size_t)N, stream));
    exclusive_sum(count.as<int>(), pat.elem_ptr.as<int>(), N + 1, stream, tmp);
    k_elem_fill<<<grid_for(E, kBlock, 65536), kBlock, 0, stream>>>(E, ne, pat.conn.as<int>(), pat.elem_ptr.as<int>(),
                                                                   cursor.as<int>(), pat.elem_list.as<uint32_t>());
    k_pei_sort<<<grid_for(N, kBlock, 65536), kBlock, 0, stream>>>(N, pat.elem_ptr.as<int>(),
                                                                  pat.elem_list.as<uint32_t>());
    const long long total = (long long)E * ne;
    pat.elem_nodes.reserve(sizeof(int4) * (size_t)total);
    k_elem_nodes<<<grid_for(total, kBlock, 65536), kBlock, 0, stream>>>(total, E, ne, pat.conn.as<int>(),
                                                                        pat.elem_list.as<uint32_t>(),
                                                                        pat.elem_nodes.as<int4>());
    PFEM_HIP(hipStreamSynchronize(stream));
}

void build_pattern(int dim, int N, int E, const int* conn_aos, const uint8_t* flags_in, hipStream_t stream,
                   Pattern& pat) {
    PFEM_REQUIRE(dim == 2 || dim == 3, E_INVALID, "dim must be 2 or 3");
    PFEM_REQUIRE(N > 0 && E > 0, E_INVALID, "mesh must have nodes and elements");
    const int ne = dim + 1;
    const long long npairs = (long long)E * ne * ne;
    PFEM_REQUIRE(npairs < (1LL << 31), E_INVALID, "too many elements for int32 pair indexing");
    PFEM_REQUIRE((long long)(dim + 1) * N < (1LL << 31), E_INVALID, "too many DOFs for int32 indexing");
    pat.dim = dim;
    pat.N = N;
    pat.E = E;
    pat.n = (dim + 1) * N;

    DevBuf count, err, keys, keys2, tmp, nsel, len_m, len_c, ptr_m, ptr_c, maxdeg;
    pat.conn.reserve(sizeof(int) * (size_t)E * ne);
    pat.flags.reserve((size_t)N);
    count.reserve(sizeof(int) * (size_t)N);
    err.reserve(sizeof(int));
    PFEM_HIP(hipMemsetAsync(count.p, 0, sizeof(int) * (size_t)N, stream));
    PFEM_HIP(hipMemsetAsync(err.p, 0, sizeof(int), stream));
    const int B = kBlock;
    k_conn_soa<<<grid_for(E, B, 65536), B, 0, stream>>>(E, ne, N, conn_aos, pat.conn.as<int>(), count.as<int>(),
                                                        err.as<int>());
    k_flags<<<grid_for(N, B, 65536), B, 0, stream>>>(N, flags_in, count.as<int>(), pat.flags.as<uint8_t>());

    keys.reserve(sizeof(unsigned long long) * npairs);
    keys2.reserve(sizeof(unsigned long long) * npairs);
    if (dim == 2)
        k_pair_keys<3><<<grid_for(E, B, 65536), B, 0, stream>>>(E, N, pat.conn.as<int>(),
                                                                keys.as<unsigned long long>());
    else
        k_pair_keys<4><<<grid_for(E, B, 65536), B, 0, stream>>>(E, N, pat.conn.as<int>(),
                                                                keys.as<unsigned long long>());
    const unsigned long long maxkey = (unsigned long long)N * (unsigned long long)N;
    int bits = 1;
    while (bits < 64 && (1ULL << bits) < maxkey) ++bits;
    size_t bytes = 0;
    PFEM_HIP(hipcub::DeviceRadixSort::SortKeys(nullptr, bytes, keys.as<unsigned long long>(),
                                               keys2.as<unsigned long long>(), (int)npairs, 0, bits, stream));
    tmp.reserve(bytes);
    PFEM_HIP(hipcub::DeviceRadixSort::SortKeys(tmp.p, bytes, keys.as<unsigned long long>(),
                                               keys2.as<unsigned long long>(), (int)npairs, 0, bits, stream));
    nsel.reserve(sizeof(int));
    bytes = 0;
    PFEM_HIP(hipcub::DeviceSelect::Unique(nullptr, bytes, keys2.as<unsigned long long>(),
                                          keys.as<unsigned long long>(), nsel.as<int>(), (int)npairs, stream));
    tmp.reserve(bytes);
    PFEM_HIP(hipcub::DeviceSelect::Unique(tmp.p, bytes, keys2.as<unsigned long long>(),
                                          keys.as<unsigned long long>(), nsel.as<int>(), (int)npairs, stream));
    int M = 0, herr = 0;
    PFEM_HIP(hipMemcpyAsync(&M, nsel.p, sizeof(int), hipMemcpyDeviceToHost, stream));
    PFEM_HIP(hipMemcpyAsync(&herr, err.p, sizeof(int), hipMemcpyDeviceToHost, stream));
    PFEM_HIP(hipStreamSynchronize(stream));
    PFEM_REQUIRE(herr == 0, E_INVALID, "connectivity references a node index outside [0, N)");
    pat.graph_nnz = M;
    const unsigned long long* ukeys = keys.as<unsigned long long>();

    pat.nbr_ptr.reserve(sizeof(int) * (size_t)(N + 1));
    pat.nbr_idx.reserve(sizeof(int) * (size_t)M);
    k_nbr_ptr<<<grid_for(N + 1, B, 65536), B, 0, stream>>>(N, M, ukeys, pat.nbr_ptr.as<int>());
    k_nbr_idx<<<grid_for(M, B, 65536), B, 0, stream>>>(M, N, ukeys, pat.nbr_idx.as<int>());

    len_m.reserve(sizeof(int) * (size_t)(N + 1));
    len_c.reserve(sizeof(int) * (size_t)(N + 1));
    ptr_m.reserve(sizeof(int) * (size_t)(N + 1));
    ptr_c.reserve(sizeof(int) * (size_t)(N + 1));
    k_row_len<<<grid_for(N + 1, B, 65536), B, 0, stream>>>(N, dim, pat.nbr_ptr.as<int>(), pat.flags.as<uint8_t>(),
                                                           len_m.as<int>(), len_c.as<int>());
    exclusive_sum(len_m.as<int>(), ptr_m.as<int>(), N + 1, stream, tmp);
    exclusive_sum(len_c.as<int>(), ptr_c.as<int>(), N + 1, stream, tmp);
    int tot[2] = {0, 0};
    PFEM_HIP(hipMemcpyAsync(&tot[0], ptr_m.as<int>() + N, sizeof(int), hipMemcpyDeviceToHost, stream));
    PFEM_HIP(hipMemcpyAsync(&tot[1], ptr_c.as<int>() + N, sizeof(int), hipMemcpyDeviceToHost, stream));
    maxdeg.reserve(sizeof(int));
    PFEM_HIP(hipMemsetAsync(maxdeg.p, 0, sizeof(int), stream));
    pat.node_info.reserve(sizeof(int4) * (size_t)N);
    k_node_info<<<grid_for(N, B, 65536), B, 0, stream>>>(N, ptr_m.as<int>(), ptr_c.as<int>(), pat.nbr_ptr.as<int>(),
                                                         pat.flags.as<uint8_t>(), pat.node_info.as<int4>(),
                                                         maxdeg.as<int>());
    int hmax = 0;
    PFEM_HIP(hipMemcpyAsync(&hmax, maxdeg.p, sizeof(int), hipMemcpyDeviceToHost, stream));
    PFEM_HIP(hipStreamSynchronize(stream));
    pat.max_deg = hmax;
    PFEM_REQUIRE(hmax <= 255, E_UNSUPPORTED, "node degree above 255 is not supported");
    pat.mom_nnz = tot[0];
    pat.cont_nnz = tot[1];
    pat.nnz = (long long)dim * tot[0] + tot[1];
    PFEM_REQUIRE(pat.nnz < (1LL << 31), E_UNSUPPORTED, "nnz above 2^31 is not supported");

    const long long n = pat.n;
    pat.row_ptr.reserve(sizeof(int) * (size_t)(n + 1));
    pat.diag_pos.reserve(sizeof(int) * (size_t)n);
    k_row_ptr<<<grid_for(n + 1, B, 65536), B, 0, stream>>>(N, dim, pat.mom_nnz, pat.nnz, pat.node_info.as<int4>(),
                                                           pat.row_ptr.as<int>());
    k_col_diag<<<grid_for(N, B, 65536), B, 0, stream>>>(N, dim, pat.mom_nnz, pat.node_info.as<int4>(),
                                                        pat.nbr_ptr.as<int>(), pat.nbr_idx.as<int>(), nullptr,
                                                        pat.diag_pos.as<int>());
    // element slot ranks: only needed to build the pair-element incidence
    DevBuf ranks;
    ranks.reserve(16 * (size_t)E);
    PFEM_HIP(hipMemsetAsync(err.p, 0, sizeof(int), stream));
    if (dim == 2)
        k_ranks<3><<<grid_for(E, B, 65536), B, 0, stream>>>(E, pat.conn.as<int>(), pat.nbr_ptr.as<int>(),
                                                            pat.nbr_idx.as<int>(), ranks.as<uint8_t>(),
                                                            err.as<int>());
    else
        k_ranks<4><<<grid_for(E, B, 65536), B, 0, stream>>>(E, pat.conn.as<int>(), pat.nbr_ptr.as<int>(),
                                                            pat.nbr_idx.as<int>(), ranks.as<uint8_t>(),
                                                            err.as<int>());
    PFEM_HIP(hipMemcpyAsync(&herr, err.p, sizeof(int), hipMemcpyDeviceToHost, stream));
    PFEM_HIP(hipStreamSynchronize(stream));
    PFEM_REQUIRE(herr == 0, E_INVALID, "element slot ranks inconsistent with the node graph");
    // pair-element incidence (gather assembly)
    PFEM_REQUIRE((long long)E < (1LL << 28), E_UNSUPPORTED, "more than 2^28 elements");
    DevBuf cnt, cursor;
    cnt.reserve(sizeof(int) * (size_t)(M + 1));
    cursor.reserve(sizeof(int) * (size_t)M);
    PFEM_HIP(hipMemsetAsync(cnt.p, 0, sizeof(int) * (size_t)(M + 1), stream));
    PFEM_HIP(hipMemsetAsync(cursor.p, 0, sizeof(int) * (size_t)M, stream));
    pat.pei_ptr.reserve(sizeof(int) * (size_t)(M + 1));
    pat.pei.reserve(sizeof(uint32_t) * (size_t)npairs);
    if (dim == 2)
        k_pei_count<3><<<grid_for(E, B, 65536), B, 0, stream>>>(E, pat.conn.as<int>(), ranks.as<uint4>(),
                                                                pat.nbr_ptr.as<int>(), cnt.as<int>());
    else
        k_pei_count<4><<<grid_for(E, B, 65536), B, 0, stream>>>(E, pat.conn.as<int>(), ranks.as<uint4>(),
                                                                pat.nbr_ptr.as<int>(), cnt.as<int>());
    exclusive_sum(cnt.as<int>(), pat.pei_ptr.as<int>(), M + 1, stream, tmp);
    if (dim == 2)
        k_pei_fill<3><<<grid_for(E, B, 65536), B, 0, stream>>>(E, pat.conn.as<int>(), ranks.as<uint4>(),
                                                               pat.nbr_ptr.as<int>(), pat.pei_ptr.as<int>(),
                                                               cursor.as<int>(), pat.pei.as<uint32_t>());
    else
        k_pei_fill<4><<<grid_for(E, B, 65536), B, 0, stream>>>(E, pat.conn.as<int>(), ranks.as<uint4>(),
                                                               pat.nbr_ptr.as<int>(), pat.pei_ptr.as<int>(),
                                                               cursor.as<int>(), pat.pei.as<uint32_t>());
    k_pei_sort<<<grid_for(M, B, 65536), B, 0, stream>>>(M, pat.pei_ptr.as<int>(), pat.pei.as<uint32_t>());
    PFEM_HIP(hipGetLastError());
    PFEM_HIP(hipStreamSynchronize(stream));
}

void build_col_idx(const Pattern& pat, int* col_idx, hipStream_t stream) {
    const int B = kBlock;
    k_col_node<<<grid_for(pat.N, B, 65536), B, 0, stream>>>(pat.N, pat.dim, pat.mom_nnz, pat.node_info.as<int4>(),
                                                            pat.nbr_ptr.as<int>(), pat.nbr_idx.as<int>(), col_idx);
    PFEM_HIP(hipGetLastError());
}

}  // namespace pfem
