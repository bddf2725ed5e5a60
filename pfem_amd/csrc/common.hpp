// Shared device helpers for the PFEM gfx950 kernels.
//
// Every kernel in this library is launched over a *work table*: a device array of
// BlockTask entries, one per workgroup.  Entry b tells workgroup b which case it works
// on and which slice of that case it owns (block index and block count within the
// case); work inside a case is grid-strided over that slice.  A single case is a table
// with one case; a serving batch is a table with many.  Because the table lives in
// device memory, the same launch (or captured hipGraph) serves any batch composition.
//
// Reductions are deterministic: each workgroup reduces in a fixed tree, writes one
// partial per quantity, and the last workgroup of the case (detected with a counter)
// sums the partials in block order and finalises the case's scalars.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

namespace pfem {

constexpr int kBlock = 256;  // threads per workgroup (4 wavefronts of 64)
constexpr int kWaves = kBlock / 64;
constexpr int kXcd = 8;  // XCDs of the chip; workgroup g of a launch runs on XCD g % kXcd

struct BlockTask {
    int case_id;  // index into the launch's case array; < 0 marks an idle workgroup
    int block;    // workgroup index within the case
    int nblocks;  // workgroups given to the case in this launch
    int pad;
};

// XCD-aware work split: workgroups are dispatched round-robin over the kXcd XCDs; the
// workgroups of a case that land on one XCD own one contiguous range of its `nodes` items,
// so neighbouring items (and the data they share) stay in that XCD's L2.
struct Seg {
    int n0;   // first node of the range
    int ns;   // nodes in the range
    int lb;   // index of this workgroup among the case's workgroups on the range
    int cnt;  // workgroups of the case on the range
};

__device__ __forceinline__ Seg xcd_segment(const BlockTask& t, int nodes) {
    int seg, nseg, lb, cnt;
    if (t.nblocks < kXcd) {
        seg = t.block;
        nseg = t.nblocks;
        lb = 0;
        cnt = 1;
    } else {
        const int g0 = (int)blockIdx.x - t.block;  // global id of the case's first workgroup
        seg = ((int)blockIdx.x % kXcd - g0 % kXcd + kXcd) % kXcd;  // == t.block % kXcd
        nseg = kXcd;
        lb = (t.block - seg) / kXcd;
        cnt = (t.nblocks - seg + kXcd - 1) / kXcd;
    }
    Seg s;
    s.n0 = (int)((long long)nodes * seg / nseg);
    s.ns = (int)((long long)nodes * (seg + 1) / nseg) - s.n0;
    s.lb = lb;
    s.cnt = cnt;
    return s;
}

// Node flag bits (PFEM_NODE_* in include/pfem_gpu.h).
enum : uint8_t { NODE_BOUND = 1, NODE_FREE = 2, NODE_ONFS = 4, NODE_FIXED = 8, NODE_DIRICHLET = 16, NODE_DIRICHLET_T = 32 };

// Lane permutations inside 16-lane DPP rows (VALU moves; no LDS round trip as __shfl_xor's
// ds_bpermute takes).  CTRL is a DPP control: 0xB1 / 0x4E quad_perm [1,0,3,2] / [2,3,0,1],
// 0x141 row_half_mirror (lane i of 8 reads 7-i), 0x140 row_mirror (lane i of 16 reads 15-i).
template <int CTRL>
__device__ __forceinline__ float dpp_mov(float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, 0xF, 0xF, false));
}
template <int CTRL>
__device__ __forceinline__ double dpp_mov(double v) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)b, CTRL, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, 0xF, 0xF, false);
    return __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
}

// Sum over each aligned group of L lanes (L = 4, 8 or 16), returned in every lane of the
// group.  The partner sums are symmetric, so all lanes of a group hold the same bits.  Every
// lane of a group must be active (groups never straddle a DPP row, so other groups may not be).
template <int L, typename T>
__device__ __forceinline__ T group_sum(T v) {
    static_assert(L == 4 || L == 8 || L == 16, "group width");
    v += dpp_mov<0xB1>(v);
    v += dpp_mov<0x4E>(v);
    if (L >= 8) v += dpp_mov<0x141>(v);
    if (L >= 16) v += dpp_mov<0x140>(v);
    return v;
}

// Smallest value over each aligned group of L lanes, in every lane of the group (same lane
// rules as group_sum).
template <int L>
__device__ __forceinline__ double group_min(double v) {
    static_assert(L == 4 || L == 8 || L == 16, "group width");
    v = fmin(v, dpp_mov<0xB1>(v));
    v = fmin(v, dpp_mov<0x4E>(v));
    if (L >= 8) v = fmin(v, dpp_mov<0x141>(v));
    if (L >= 16) v = fmin(v, dpp_mov<0x140>(v));
    return v;
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
    return v;
}

__device__ __forceinline__ double wave_min(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmin(v, __shfl_xor(v, o, 64));
    return v;
}

// Gauss-Jordan inversion of a 4 x 4 block with partial pivoting on [M | I] (m[r][0..3] = M,
// m[r][4..7] = I on entry; the inverse in m[r][4..7] on exit).  Fully unrolled, the pivot
// row swap done with selects, so the block stays in registers (a swap indexed by the
// run-time pivot row put it in scratch memory).  ZERO_SAFE: a zero pivot gives a zero row
// scale instead of an infinity.
template <bool ZERO_SAFE>
__device__ __forceinline__ void gj_invert4(double (&m)[4][8]) {
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        int piv = c;
        double best = fabs(m[c][c]);
#pragma unroll
        for (int r = c + 1; r < 4; ++r) {
            const double a = fabs(m[r][c]);
            if (a > best) {
                best = a;
                piv = r;
            }
        }
#pragma unroll
        for (int r = c + 1; r < 4; ++r) {
            const bool sw = (r == piv);
#pragma unroll
            for (int k = 0; k < 8; ++k) {
                const double a = m[c][k], b = m[r][k];
                m[c][k] = sw ? b : a;
                m[r][k] = sw ? a : b;
            }
        }
        const double inv = ZERO_SAFE ? (m[c][c] != 0.0 ? 1.0 / m[c][c] : 0.0) : 1.0 / m[c][c];
#pragma unroll
        for (int k = 0; k < 8; ++k) m[c][k] *= inv;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            if (r == c) continue;
            const double fct = m[r][c];
            if (fct != 0.0)
#pragma unroll
                for (int k = 0; k < 8; ++k) m[r][k] -= fct * m[c][k];
        }
    }
}

// Smallest positive value over the workgroup folded into *addr (the bits of a positive
// double, compared as integers): one atomic per workgroup, since every atomic on one address
// is serialised (one per wavefront cost ~10 ns each, 0.3 ms over 2M nodes).
__device__ __forceinline__ void block_min_to(unsigned long long* addr, double v) {
    __shared__ double red[kWaves];
    v = wave_min(v);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double m = red[0];
#pragma unroll
        for (int w = 1; w < kWaves; ++w) m = fmin(m, red[w]);
        atomicMin(addr, (unsigned long long)__double_as_longlong(m));
    }
}

// Sum NV values over the workgroup; the result is valid in thread 0.
template <int NV>
__device__ __forceinline__ void block_sum(double (&v)[NV]) {
    __shared__ double red[NV][kWaves];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int j = 0; j < NV; ++j) {
        double s = wave_sum(v[j]);
        if (lane == 0) red[j][wave] = s;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int j = 0; j < NV; ++j) {
            double s = red[j][0];
#pragma unroll
            for (int w = 1; w < kWaves; ++w) s += red[j][w];
            v[j] = s;
        }
    }
    __syncthreads();
}

// Two-level deterministic reduction.  Every workgroup of a case calls this with its
// thread-local contributions; nblocks must not exceed kMaxCaseBlocks (views.hpp), the
// size of each case's partials buffer in NV-tuples.  Returns true in exactly one workgroup (the last to
// arrive), in which `tot` (valid in thread 0) holds the case-wide sums taken in block
// order.  `partials` holds nblocks*NV doubles for this case; `counter` must be zero on
// entry and is reset to zero by the last workgroup.
template <int NV>
__device__ bool reduce_last_block(double (&v)[NV], double* partials, unsigned* counter, int block, int nblocks,
                                  double (&tot)[NV]) {
    __shared__ bool is_last;
    block_sum<NV>(v);
    if (threadIdx.x == 0) {
#pragma unroll
        for (int j = 0; j < NV; ++j)
            __hip_atomic_store(&partials[block * NV + j], v[j], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __threadfence();
        unsigned prev = atomicAdd(counter, 1u);
        is_last = (prev == (unsigned)(nblocks - 1));
    }
    __syncthreads();
    if (!is_last) return false;
    __threadfence();
    double acc[NV];
#pragma unroll
    for (int j = 0; j < NV; ++j) acc[j] = 0.0;
    // thread t sums blocks t, t + kBlock, ... in increasing order; then a fixed tree
    for (int bidx = threadIdx.x; bidx < nblocks; bidx += kBlock) {
#pragma unroll
        for (int j = 0; j < NV; ++j)
            acc[j] += __hip_atomic_load(&partials[bidx * NV + j], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    block_sum<NV>(acc);
    if (threadIdx.x == 0) {
#pragma unroll
        for (int j = 0; j < NV; ++j) tot[j] = acc[j];
        __hip_atomic_store(counter, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    return true;
}

// Positive doubles order like their bit patterns: min over positive values via integer atomics.
__device__ __forceinline__ void atomic_min_pos_double(unsigned long long* addr, double v) {
    atomicMin(addr, (unsigned long long)__double_as_longlong(v));
}

// Load of data read once per kernel (matrix values of an SpMV): the non-temporal hint keeps
// the stream from evicting the gathered vector and the stored results from the caches.
// Measured on gfx950 (ELL SpMV, 516 MB matrix): 98.7 -> 90.0 us per launch.
template <typename T>
__device__ __forceinline__ T ld_stream(const T* p) {
    return __builtin_nontemporal_load(p);
}

}  // namespace pfem
