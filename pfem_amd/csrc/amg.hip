// Aggregation multigrid for the PSPG systems: per-assembly setup (Galerkin coarse
// operators, block-Jacobi inverses, dense inverse of the coarsest level) and the V-cycle
// used as BiCGSTAB's right preconditioner, each batched over cases and levels through work
// tables.  The hierarchy's structure comes from amg_hierarchy.cpp; see amg.hpp for the
// transfer.
//
// V-cycle on level l (input r_l, output z_l; level 0 reads p or s and writes ph or sh):
//   w = omega B^-1 r                      (k_amg_pre at level 0, block-Jacobi pre-smoothing;
//                                          on coarser levels done by the restriction above)
//   t = r - A w                           (k_amg_res)
//   r_{l+1} = P^T t                       (k_amg_restrict, gather over aggregate members)
//   z_{l+1} = V-cycle(l+1)  /  A_c^-1 r_c (k_amg_coarsest, dense inverse on the coarsest)
//   w' = w + alpha P z_{l+1}              (alpha > 1 over-corrects the piecewise-constant
//                                          interpolation)
//   z = w' + omega B^-1 (r - A w')        (k_amg_post, post-smoothing; r - A w' is formed
//       = w' + omega B^-1 (t - alpha (A P) z_{l+1})   from t with the set-up product A P)
// Every reduction is a fixed-order gather, so a case's V-cycle does not depend on the batch.
#include <type_traits>

#include "common.hpp"
#include "ellspmv.hpp"
#include "nodespmv.hpp"
#include "views.hpp"

namespace pfem {

namespace {

__device__ __forceinline__ bool row_full_flags(uint8_t f, int s, int D) {
    return (s < D) ? (f & (NODE_BOUND | NODE_FREE)) == 0 : (f & NODE_FREE) == 0;
}

__device__ __forceinline__ long long row_start_of(const int4& inf, int s, int D, int mom_nnz) {
    return (s < D) ? (long long)s * mom_nnz + inf.x : (long long)D * mom_nnz + inf.y;
}

// Entry q of the level's input (level 0: the FP64 Krylov vector p or s).
template <typename T>
__device__ __forceinline__ T in_at(const AmgView<T>& V, int which, long long q) {
    return V.level == 0 ? (T)V.r0[which][q] : V.r[q];
}

}  // namespace

// ---- setup --------------------------------------------------------------------------------

// Galerkin coarse operator of level l >= 1: entry block (I, J) = P^T A P = the sum over the
// members i of aggregate I of the finer level's A P block (i, J), gathered in member order
// (A P already holds the finer blocks (i, j) summed over j in J, level 0's masked to full rows
// and columns).  L lanes per coarse node, lane k owns entry k of the node's list.
template <typename T, typename ST, int L>
__device__ __forceinline__ void galerkin_body(const AmgView<T>& V, const EllMatT<ST>& F, const BlockTask& task) {
    constexpr int NPB = kBlock / L;
    const int lane = threadIdx.x % L;
    const int nf = V.A.nf, D = nf - 1, nf2 = nf * nf;
    for (long long I = (long long)task.block * NPB + threadIdx.x / L; I < V.A.N; I += (long long)task.nblocks * NPB) {
        const int4 inf = V.A.info[I];
        const int deg = inf.z;
        const int self = V.self_rank[I];
        for (int k = lane; k < deg; k += L) {
            const int mc = V.A.nbr_ptr[I] + k;
            double blk[kMaxNf][kMaxNf];
#pragma unroll
            for (int s = 0; s < kMaxNf; ++s)
#pragma unroll
                for (int t = 0; t < kMaxNf; ++t) blk[s][t] = 0.0;
            for (int q = V.gal_ptr[mc]; q < V.gal_ptr[mc + 1]; ++q) {
                const int2 src = V.gal_src[q];
                const long long c = src.x / kEllSlice, r = src.x % kEllSlice;
                const ST* b = F.vals + ((long long)F.soff[c] + src.y) * nf2 * kEllSlice + r;
#pragma unroll
                for (int s = 0; s < kMaxNf; ++s) {
                    if (s >= nf) continue;
#pragma unroll
                    for (int t = 0; t < kMaxNf; ++t)
                        if (t < nf) blk[s][t] += (double)b[(long long)(s * nf + t) * kEllSlice];
                }
            }
#pragma unroll
            for (int s = 0; s < kMaxNf; ++s) {
                if (s >= nf) continue;
                const long long rs = row_start_of(inf, s, D, V.A.mom_nnz) + k;
#pragma unroll
                for (int t = 0; t < kMaxNf; ++t) {
                    if (t >= nf) continue;
                    double v = blk[s][t];
                    if (k == self && s == t && v == 0.0) v = 1.0;  // DOF with no member DOF: decoupled
                    V.vals_out[rs + (long long)t * deg] = (T)v;
                }
            }
        }
    }
}

template <typename T, typename ST>
__device__ __forceinline__ void galerkin_lanes(const AmgView<T>& V, const EllMatT<ST>& F, const BlockTask& task) {
    if (V.lanes == 4) galerkin_body<T, ST, 4>(V, F, task);
    else if (V.lanes == 8) galerkin_body<T, ST, 8>(V, F, task);
    else galerkin_body<T, ST, 16>(V, F, task);
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_amg_galerkin(const AmgView<T>* __restrict__ views,
                                                         const BlockTask* __restrict__ tasks) {
    const BlockTask task = tasks[blockIdx.x];
    if (task.case_id < 0) return;
    const AmgView<T>& V = views[task.case_id];
    if (V.FAPf.vals) galerkin_lanes<T, float>(V, V.FAPf, task);
    else galerkin_lanes<T, T>(V, V.FAP, task);
}

// Block-Jacobi inverse of a coarse level (the diagonal nf x nf block of each node).
template <typename T>
__global__ __launch_bounds__(kBlock) void k_amg_bjac(const AmgView<T>* __restrict__ views,
                                                     const BlockTask* __restrict__ tasks) {
    const BlockTask task = tasks[blockIdx.x];
    if (task.case_id < 0) return;
    const AmgView<T>& V = views[task.case_id];
    const int nf = V.A.nf, D = nf - 1, N = V.A.N;
    for (long long I = (long long)task.block * kBlock + threadIdx.x; I < N; I += (long long)task.nblocks * kBlock) {
        const int4 inf = V.A.info[I];
        const int self = V.self_rank[I];
        double m[4][8];
#pragma unroll
        for (int s = 0; s < 4; ++s)
#pragma unroll
            for (int t = 0; t < 8; ++t) m[s][t] = (t - 4 == s) ? 1.0 : 0.0;
#pragma unroll
        for (int s = 0; s < 4; ++s) {
            if (s >= nf) {
                m[s][s] = 1.0;  // padding rows: identity
                continue;
            }
            const long long rs = row_start_of(inf, s, D, V.A.mom_nnz) + self;
#pragma unroll
            for (int t = 0; t < 4; ++t)
                if (t < nf) m[s][t] = (double)V.A.vals[rs + (long long)t * inf.z];
        }
        gj_invert4<true>(m);
#pragma unroll
        for (int s = 0; s < 4; ++s)
#pragma unroll
            for (int t = 0; t < 4; ++t)
                if (s < nf && t < nf) V.binv_out[(long long)(s * nf + t) * N + I] = (T)m[s][4 + t];
    }
}

// A P of a level with a coarser one: entry block (i, J) sums the blocks (i, j) of i's
// neighbours j in aggregate J (gather list, fixed order).  Level 0 keeps only full rows
// (i, s) and full columns (j, t): P prolongates to full rows only, and a row holding only
// its diagonal has no entry off it.  Written in the sliced ELL layout, padding ranks of
// each node's slice as zeros.
template <typename T, int L>
__device__ __forceinline__ void ap_body(const AmgView<T>& V, const BlockTask& task) {
    constexpr int NPB = kBlock / L;
    const int lane = threadIdx.x % L;
    const int nf = V.A.nf, D = nf - 1;
    const NodeMat<T>& A = V.A;
    const EllMatT<T>& P = V.AP;
    for (long long i = (long long)task.block * NPB + threadIdx.x / L; i < P.N; i += (long long)task.nblocks * NPB) {
        const int e0 = V.ap_ptr[i], pdeg = V.ap_ptr[i + 1] - e0;
        const long long c = i / kEllSlice, r = i % kEllSlice;
        const long long b = P.soff[c];
        const int W = P.soff[c + 1] - (int)b;
        const int4 ainf = A.info[i];
        const int anb = A.nbr_ptr[i];
        const uint8_t fi = (uint8_t)ainf.w;
        for (int m = lane; m < W; m += L) {
            double blk[kMaxNf][kMaxNf];
#pragma unroll
            for (int s = 0; s < kMaxNf; ++s)
#pragma unroll
                for (int t = 0; t < kMaxNf; ++t) blk[s][t] = 0.0;
            const int qe = m < pdeg ? V.ap_gptr[e0 + m + 1] : 0;
            for (int q = m < pdeg ? V.ap_gptr[e0 + m] : 0; q < qe; ++q) {
                const int k = V.ap_gsrc[q];
                const uint8_t fj = V.mask ? V.flags[A.nbr_idx[anb + k]] : (uint8_t)0;
#pragma unroll
                for (int s = 0; s < kMaxNf; ++s) {
                    if (s >= nf) continue;
                    if (V.mask && !row_full_flags(fi, s, D)) continue;
                    const long long rs = row_start_of(ainf, s, D, A.mom_nnz) + k;
#pragma unroll
                    for (int t = 0; t < kMaxNf; ++t) {
                        if (t >= nf) continue;
                        if (V.mask && !row_full_flags(fj, t, D)) continue;
                        blk[s][t] += (double)A.vals[rs + (long long)t * ainf.z];
                    }
                }
            }
#pragma unroll
            for (int s = 0; s < kMaxNf; ++s)
#pragma unroll
                for (int t = 0; t < kMaxNf; ++t)
                    if (s < nf && t < nf) {
                        const long long q = ((b + m) * nf * nf + s * nf + t) * kEllSlice + r;
                        if (V.ap_out32) V.ap_out32[q] = (float)blk[s][t];
                        else V.ap_out[q] = (T)blk[s][t];
                    }
        }
    }
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_amg_ap(const AmgView<T>* __restrict__ views,
                                                   const BlockTask* __restrict__ tasks) {
    const BlockTask task = tasks[blockIdx.x];
    if (task.case_id < 0) return;
    const AmgView<T>& V = views[task.case_id];
    if (V.ap_lanes == 4) ap_body<T, 4>(V, task);
    else if (V.ap_lanes == 8) ap_body<T, 8>(V, task);
    else ap_body<T, 16>(V, task);
}

// Dense inverse of the coarsest level (one workgroup per case; Gauss-Jordan in LDS without
// pivoting: the coarse Galerkin operators keep positive diagonals; a vanishing pivot is
// replaced by 1, which decouples that unknown).
template <typename T>
__global__ __launch_bounds__(kBlock) void k_amg_dense(const AmgView<T>* __restrict__ views,
                                                      const BlockTask* __restrict__ tasks) {
    __shared__ double a[kAmgMaxDense * kAmgMaxDense];
    __shared__ double pivrow[kAmgMaxDense];
    __shared__ double colk[kAmgMaxDense];
    const BlockTask task = tasks[blockIdx.x];
    if (task.case_id < 0) return;
    const AmgView<T>& V = views[task.case_id];
    const int n = V.nc, N = V.A.N, nf = V.A.nf, D = nf - 1;
    for (int e = threadIdx.x; e < n * n; e += kBlock) a[e] = 0.0;
    __syncthreads();
    // scatter the node-block values: unknown (I, s) is I + s N
    for (int I = threadIdx.x; I < N; I += kBlock) {
        const int4 inf = V.A.info[I];
        const int nb = V.A.nbr_ptr[I];
        for (int k = 0; k < inf.z; ++k) {
            const int J = V.A.nbr_idx[nb + k];
            for (int s = 0; s < nf; ++s) {
                const long long rs = row_start_of(inf, s, D, V.A.mom_nnz) + k;
                for (int t = 0; t < nf; ++t) a[(I + s * N) * n + (J + t * N)] = (double)V.A.vals[rs + (long long)t * inf.z];
            }
        }
    }
    __syncthreads();
    // in-place Gauss-Jordan inversion with partial pivoting: step k brings the largest entry of
    // column k (rows >= k) to the diagonal, scales row k by its inverse and eliminates column k
    // from every other row, storing the inverse's column k in place of the eliminated one; the
    // row exchanges are undone on the inverse's columns at the end, last one first.  A column
    // that is zero on and below the diagonal (an exactly singular operator) keeps a unit pivot.
    __shared__ int perm[kAmgMaxDense];
    for (int k = 0; k < n; ++k) {
        if (threadIdx.x == 0) {
            int r = k;
            double best = fabs(a[k * n + k]);
            for (int i = k + 1; i < n; ++i) {
                const double v = fabs(a[i * n + k]);
                if (v > best) {
                    best = v;
                    r = i;
                }
            }
            perm[k] = r;
        }
        __syncthreads();
        const int r = perm[k];
        if (r != k)
            for (int j = threadIdx.x; j < n; j += kBlock) {
                const double t = a[k * n + j];
                a[k * n + j] = a[r * n + j];
                a[r * n + j] = t;
            }
        __syncthreads();
        double p = a[k * n + k];
        p = (p != 0.0) ? 1.0 / p : 1.0;
        for (int j = threadIdx.x; j < n; j += kBlock) {
            pivrow[j] = (j == k) ? p : a[k * n + j] * p;
            colk[j] = a[j * n + k];
        }
        __syncthreads();
        for (int e = threadIdx.x; e < n * n; e += kBlock) {
            const int i = e / n, j = e - i * n;
            if (i == k) a[e] = pivrow[j];
            else if (j == k) a[e] = -colk[i] * p;
            else a[e] -= colk[i] * pivrow[j];
        }
        __syncthreads();
    }
    for (int k = n - 1; k >= 0; --k) {
        const int r = perm[k];
        if (r != k)
            for (int i = threadIdx.x; i < n; i += kBlock) {
                const double t = a[i * n + k];
                a[i * n + k] = a[i * n + r];
                a[i * n + r] = t;
            }
        __syncthreads();
    }
    for (int e = threadIdx.x; e < n * n; e += kBlock) V.dense[e] = a[e];
}

// ---- V-cycle -----------------------------------------------------------------------------

template <typename T>
__global__ __launch_bounds__(kBlock) void k_amg_pre(const AmgView<T>* __restrict__ views,
                                                    const BlockTask* __restrict__ tasks, int which) {
    const BlockTask task = tasks[blockIdx.x];
    if (task.case_id < 0) return;
    const AmgView<T>& V = views[task.case_id];
    if (*V.done_flag) return;
    const int nf = V.A.nf, N = V.A.N;
    const T om = (T)V.omega;
    for (long long i = (long long)task.block * kBlock + threadIdx.x; i < N; i += (long long)task.nblocks * kBlock) {
        T x[kMaxNf];
#pragma unroll
        for (int t = 0; t < kMaxNf; ++t) x[t] = (t < nf) ? in_at(V, which, i + (long long)t * N) : T(0);
#pragma unroll
        for (int s = 0; s < kMaxNf; ++s) {
            if (s >= nf) continue;
            T acc = 0;
#pragma unroll
            for (int t = 0; t < kMaxNf; ++t)
                if (t < nf) acc += V.binv[(long long)(s * nf + t) * N + i] * x[t];
            V.w[i + (long long)s * N] = om * acc;
        }
    }
}

template <typename T, int L>
__device__ __forceinline__ void res_body(const AmgView<T>& V, const BlockTask& task, int which) {
    constexpr int NPB = kBlock / L;
    const int lane = threadIdx.x % L;
    const int N = V.A.N;
    for (long long i = (long long)task.block * NPB + threadIdx.x / L; i < N; i += (long long)task.nblocks * NPB) {
        T y[kMaxNf];
        node_spmv<L>(V.A, (const T*)V.w, i, lane, y);
        if (lane < V.A.nf) {
            const long long q = i + (long long)lane * N;
            V.t[q] = in_at(V, which, q) - pick(y, lane);
        }
    }
}

// Level 0 with the system's ELL copy: one lane per node, whole 64-node slices per
// wavefront (a workgroup's waves stride over the case's slices); an FP64 hierarchy reads
// the FP32 copy of the values when the view has one (E0f).
template <int NF, typename T>
__device__ __forceinline__ void res_ell(const AmgView<T>& V, const BlockTask& task, int which) {
    const EllMatT<T>& E = V.E0;
    const int lane = threadIdx.x & (kEllSlice - 1);
    const long long N = E.N;
    for (long long c = (long long)task.block * kWaves + (threadIdx.x >> 6); c < E.nsl;
         c += (long long)task.nblocks * kWaves) {
        T y[kMaxNf];
        if (V.E0f.vals) ell_rows<NF>(V.E0f, (const T*)V.w, c, lane, y);
        else ell_rows<NF>(E, (const T*)V.w, c, lane, y);
        const long long i = c * kEllSlice + lane;
        if (i < N) {
#pragma unroll
            for (int s = 0; s < NF; ++s) {
                const long long q = i + (long long)s * N;
                V.t[q] = (T)V.r0[which][q] - y[s];
            }
        }
    }
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_amg_res(const AmgView<T>* __restrict__ views,
                                                    const BlockTask* __restrict__ tasks, int which) {
    const BlockTask task = tasks[blockIdx.x];
    if (task.case_id < 0) return;
    const AmgView<T>& V = views[task.case_id];
    if (*V.done_flag) return;
    if (V.level == 0 && V.E0.vals) {
        if (V.A.nf == 4) res_ell<4>(V, task, which);
        else res_ell<3>(V, task, which);
        return;
    }
    if (V.lanes == 4) res_body<T, 4>(V, task, which);
    else if (V.lanes == 8) res_body<T, 8>(V, task, which);
    else res_body<T, 16>(V, task, which);
}

// r_{l+1}(I, s) = sum over the members i of aggregate I of t(i, s) (level 0: full rows only);
// kRestrictLanes lanes per aggregate share its members (aggregates hold 5-30 nodes).  When
// level l+1 is not the coarsest, its pre-smoothing w_{l+1}(I) = omega B^-1(I) r_{l+1}(I) is
// local to I and is applied here as well (level l+1 has no k_amg_pre launch).
template <typename T>
__global__ __launch_bounds__(kBlock) void k_amg_restrict(const AmgView<T>* __restrict__ views,
                                                         const BlockTask* __restrict__ tasks) {
    constexpr int L = kRestrictLanes, NPB = kBlock / L;
    const BlockTask task = tasks[blockIdx.x];
    if (task.case_id < 0) return;
    const AmgView<T>& V = views[task.case_id];
    if (*V.done_flag) return;
    const AmgView<T>& Vn = views[task.case_id + 1];  // level l+1 (views of a case are consecutive)
    const bool smooth = Vn.dense == nullptr;
    const T om = (T)Vn.omega;
    const int nf = V.A.nf, D = nf - 1, N = V.A.N, Nc = V.next_N;
    const int lane = threadIdx.x % L;
    for (long long I = (long long)task.block * NPB + threadIdx.x / L; I < Nc; I += (long long)task.nblocks * NPB) {
        double acc[kMaxNf] = {0.0, 0.0, 0.0, 0.0};
        for (int m = V.agg_ptr[I] + lane; m < V.agg_ptr[I + 1]; m += L) {
            const int i = V.agg_nodes[m];
            const uint8_t f = V.mask ? V.flags[i] : (uint8_t)0;
#pragma unroll
            for (int s = 0; s < kMaxNf; ++s)
                if (s < nf && (!V.mask || row_full_flags(f, s, D))) acc[s] += (double)V.t[i + (long long)s * N];
        }
#pragma unroll
        for (int s = 0; s < kMaxNf; ++s)
            if (s < nf) acc[s] = group_sum<L>(acc[s]);
        if (lane < nf) {
            V.next_r[I + (long long)lane * Nc] = (T)pick(acc, lane);
            if (smooth) {
                T w = 0;
#pragma unroll
                for (int t = 0; t < kMaxNf; ++t)
                    if (t < nf) w += Vn.binv[(long long)(lane * nf + t) * Nc + I] * (T)acc[t];
                Vn.w[I + (long long)lane * Nc] = om * w;
            }
        }
    }
}

// z = A_c^-1 r on the coarsest level (one workgroup per case).
template <typename T>
__global__ __launch_bounds__(kBlock) void k_amg_coarsest(const AmgView<T>* __restrict__ views,
                                                         const BlockTask* __restrict__ tasks) {
    const BlockTask task = tasks[blockIdx.x];
    if (task.case_id < 0) return;
    const AmgView<T>& V = views[task.case_id];
    if (*V.done_flag) return;
    const int n = V.nc;
    for (int row = threadIdx.x; row < n; row += kBlock) {
        double acc = 0.0;
        for (int j = 0; j < n; ++j) acc += V.dense[row * n + j] * (double)V.r[j];
        V.z[row] = (T)acc;
    }
}

// z = w' + omega B^-1 (t - alpha (A P) z_{l+1}),  w' = w + alpha P z_{l+1} (level 0: full
// rows only): one lane per node, whole 64-node slices per wavefront; A P in FP32 when the
// view holds it so (APf).
template <int NF, typename T>
__device__ __forceinline__ void post_ell(const AmgView<T>& V, const BlockTask& task, int which) {
    const int N = V.A.N, D = NF - 1, Nc = V.next_N;
    const T om = (T)V.omega, al = (T)V.alpha;
    const int lane = threadIdx.x & (kEllSlice - 1);
    for (long long c = (long long)task.block * kWaves + (threadIdx.x >> 6); c < V.AP.nsl;
         c += (long long)task.nblocks * kWaves) {
        T y[kMaxNf];
        if (V.APf.vals) ell_rows<NF>(V.APf, V.next_z, c, lane, y);
        else ell_rows<NF>(V.AP, V.next_z, c, lane, y);
        const long long i = c * kEllSlice + lane;
        if (i >= N) continue;
        T d[NF];
#pragma unroll
        for (int t = 0; t < NF; ++t) d[t] = V.t[i + (long long)t * N] - al * y[t];
        const int I = V.fine_agg[i];
        const uint8_t f = V.mask ? V.flags[i] : (uint8_t)0;
#pragma unroll
        for (int s = 0; s < NF; ++s) {
            T acc = 0;
#pragma unroll
            for (int t = 0; t < NF; ++t) acc += V.binv[(long long)(s * NF + t) * N + i] * d[t];
            const long long q = i + (long long)s * N;
            T wq = V.w[q];
            if (I >= 0 && (!V.mask || row_full_flags(f, s, D))) wq += al * V.next_z[I + (long long)s * Nc];
            const T out = wq + om * acc;
            if (V.level == 0) V.out0[which][q] = (double)out;
            else V.z[q] = out;
        }
    }
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_amg_post(const AmgView<T>* __restrict__ views,
                                                     const BlockTask* __restrict__ tasks, int which) {
    const BlockTask task = tasks[blockIdx.x];
    if (task.case_id < 0) return;
    const AmgView<T>& V = views[task.case_id];
    if (*V.done_flag) return;
    if (V.A.nf == 4) post_ell<4>(V, task, which);
    else post_ell<3>(V, task, which);
}

// ---- launchers ---------------------------------------------------------------------------
template <typename T>
void launch_amg_galerkin(const AmgView<T>* v, const BlockTask* t, int nt, hipStream_t s) {
    k_amg_galerkin<T><<<nt, kBlock, 0, s>>>(v, t);
}
template <typename T>
void launch_amg_bjac(const AmgView<T>* v, const BlockTask* t, int nt, hipStream_t s) {
    k_amg_bjac<T><<<nt, kBlock, 0, s>>>(v, t);
}
template <typename T>
void launch_amg_dense(const AmgView<T>* v, const BlockTask* t, int nt, hipStream_t s) {
    k_amg_dense<T><<<nt, kBlock, 0, s>>>(v, t);
}
template <typename T>
void launch_amg_pre(const AmgView<T>* v, const BlockTask* t, int nt, int which, hipStream_t s) {
    k_amg_pre<T><<<nt, kBlock, 0, s>>>(v, t, which);
}
template <typename T>
void launch_amg_res(const AmgView<T>* v, const BlockTask* t, int nt, int which, hipStream_t s) {
    k_amg_res<T><<<nt, kBlock, 0, s>>>(v, t, which);
}
template <typename T>
void launch_amg_restrict(const AmgView<T>* v, const BlockTask* t, int nt, hipStream_t s) {
    k_amg_restrict<T><<<nt, kBlock, 0, s>>>(v, t);
}
template <typename T>
void launch_amg_coarsest(const AmgView<T>* v, const BlockTask* t, int nt, hipStream_t s) {
    k_amg_coarsest<T><<<nt, kBlock, 0, s>>>(v, t);
}
template <typename T>
void launch_amg_ap(const AmgView<T>* v, const BlockTask* t, int nt, hipStream_t s) {
    k_amg_ap<T><<<nt, kBlock, 0, s>>>(v, t);
}
template <typename T>
void launch_amg_post(const AmgView<T>* v, const BlockTask* t, int nt, int which, hipStream_t s) {
    k_amg_post<T><<<nt, kBlock, 0, s>>>(v, t, which);
}

#define PFEM_AMG_INSTANTIATE(T)                                                                    \
    template void launch_amg_galerkin<T>(const AmgView<T>*, const BlockTask*, int, hipStream_t);   \
    template void launch_amg_bjac<T>(const AmgView<T>*, const BlockTask*, int, hipStream_t);       \
    template void launch_amg_dense<T>(const AmgView<T>*, const BlockTask*, int, hipStream_t);      \
    template void launch_amg_pre<T>(const AmgView<T>*, const BlockTask*, int, int, hipStream_t);   \
    template void launch_amg_res<T>(const AmgView<T>*, const BlockTask*, int, int, hipStream_t);   \
    template void launch_amg_restrict<T>(const AmgView<T>*, const BlockTask*, int, hipStream_t);   \
    template void launch_amg_coarsest<T>(const AmgView<T>*, const BlockTask*, int, hipStream_t);   \
    template void launch_amg_ap<T>(const AmgView<T>*, const BlockTask*, int, hipStream_t);         \
    template void launch_amg_post<T>(const AmgView<T>*, const BlockTask*, int, int, hipStream_t);
PFEM_AMG_INSTANTIATE(double)
PFEM_AMG_INSTANTIATE(float)

}  // namespace pfem
