// Node-block SpMV and right-preconditioned BiCGSTAB for a batch of cases, plus the FP64
// iterative-refinement kernels of the mixed-precision solve.
//
// Replaces SparseLU<SparseMatrix<double>, COLAMDOrdering<int>> (analyzePattern / factorize /
// solve at MomContEquationPSPG.inl:281-290) with right-preconditioned BiCGSTAB
// (van der Vorst 1992).  The preconditioner is block Jacobi over the nf x nf diagonal
// block of each node (its velocity and pressure DOFs).  One iteration is five launches:
//   K_p : p  = r + beta (p - omega v);  ph = B^-1 p        (beta from (rh, r) and slot q)
//   K_v : v  = A ph                          -> (rh, v)
//   K_s : s  = r - alpha v;  sh = B^-1 s     -> (s, s)      (alpha = rho / (rh, v))
//   K_t : t  = A sh                          -> (t, s), (t, t)   (skipped when ||s|| <= tol)
//   K_x : x += alpha ph + omega sh;  r = s - omega t -> (rh, r), (r, r)
// Reductions: every workgroup writes its FP64 partial sums to its own slot and the next
// kernel's workgroups each re-reduce them in block order — deterministic, and ordered by
// the kernel boundary, so no device-scope fence (an L2 write-back on gfx950) is needed.
// Scalars: two slots used alternately by iteration parity (KScal in views.hpp).
//
// Work split: workgroups are dispatched round-robin over the 8 XCDs; the workgroups of a
// case that land on one XCD own one contiguous range of nodes (every field of those
// nodes), so the SpMV's x gathers (neighbour nodes) and the vectors handed from one
// kernel to the next stay in that XCD's L2.  Vector kernels are node-major (a thread owns
// the nf DOFs of a node); SpMV kernels use a group of L lanes per node (nodespmv.hpp).
#include <hip/hip_ext.h>

#include <type_traits>

#include "common.hpp"
#include "ellspmv.hpp"
#include "nodespmv.hpp"
#include "views.hpp"

namespace pfem {

// Local row index l in [0, nf*ns) -> global row (field-blocked layout).
__device__ __forceinline__ long long seg_row(const Seg& s, int nodes, long long l) {
    const long long f = l / s.ns;
    return f * nodes + s.n0 + (l - f * s.ns);
}

// out = B^-1 in for one node (nf <= 4 fields, SoA blocks).
template <typename T>
__device__ __forceinline__ void block_apply(const T* __restrict__ binv, int nodes, int nf, long long i, const T (&in)[kMaxNf],
                                            T (&out)[kMaxNf]) {
#pragma unroll
    for (int s = 0; s < kMaxNf; ++s) {
        T acc = 0;
        if (s < nf) {
#pragma unroll
            for (int t = 0; t < kMaxNf; ++t)
                if (t < nf) acc += binv[(long long)(s * nf + t) * nodes + i] * in[t];
        }
        out[s] = acc;
    }
}

// The multigrid's level-0 pre-smoothing w = om B^-1 in of node i (FP64 hierarchies), written
// by the kernel that produced `in` (the same products and order as the V-cycle's own pass).
template <typename T>
__device__ __forceinline__ void smooth_w0(const KrylovView<T>& V, long long i, const T (&in)[kMaxNf]) {
    T out[kMaxNf];
    block_apply<T>(V.binv, V.nodes, V.nf, i, in, out);
    const T om = (T)V.mg_om;
#pragma unroll
    for (int s = 0; s < kMaxNf; ++s)
        if (s < V.nf) V.mg_w0[(long long)s * V.nodes + i] = (double)(om * out[s]);
}

// Sum of partial slots over the producer's workgroups, in block order; valid in every thread.
template <int NV>
__device__ __forceinline__ void reduce_partials(const double* __restrict__ P, int nblk, const int (&slot)[NV],
                                                double (&out)[NV]) {
    __shared__ double bc[NV];
    double acc[NV];
#pragma unroll
    for (int j = 0; j < NV; ++j) acc[j] = 0.0;
    for (int b = threadIdx.x; b < nblk; b += kBlock)
#pragma unroll
        for (int j = 0; j < NV; ++j) acc[j] += P[b * kPartStride + slot[j]];
    block_sum<NV>(acc);
    if (threadIdx.x == 0)
#pragma unroll
        for (int j = 0; j < NV; ++j) bc[j] = acc[j];
    __syncthreads();
#pragma unroll
    for (int j = 0; j < NV; ++j) out[j] = bc[j];
}

// Writes this workgroup's partial sums (thread-local contributions) to its slots.
template <int NV>
__device__ __forceinline__ void write_partials(double (&v)[NV], double* __restrict__ P, int block,
                                               const int (&slot)[NV]) {
    block_sum<NV>(v);
    if (threadIdx.x == 0)
#pragma unroll
        for (int j = 0; j < NV; ++j) P[block * kPartStride + slot[j]] = v[j];
}

// y = A x (plain node-block SpMV; pfem_spmv and the Picard paths).
template <typename T, int K>
__global__ __launch_bounds__(kBlock) void k_spmv(NodeMat<T> A, const T* __restrict__ x, T* __restrict__ y) {
    constexpr int NPB = kBlock / K;
    const int lane = threadIdx.x % K;
    for (long long i = (long long)blockIdx.x * NPB + threadIdx.x / K; i < A.N; i += (long long)gridDim.x * NPB) {
        T yv[kMaxNf];
        node_spmv<K>(A, x, i, lane, yv);
        if (lane < A.nf) y[i + (long long)lane * A.N] = pick(yv, lane);
    }
}

// (Re)start: r = b - A x (or b when x_zero), rh = r, p = v = 0.
template <typename T, int L>
__global__ __launch_bounds__(kBlock) void k_kry_init(const KrylovView<T>* __restrict__ views,
                                                     const BlockTask* __restrict__ tasks) {
    const BlockTask task = tasks[blockIdx.x];
    if (task.case_id < 0) return;
    const KrylovView<T>& V = views[task.case_id];
    if (!V.sc->req_init) return;
    constexpr int NPB = kBlock / L;
    const Seg sg = xcd_segment(task, V.nodes);
    const int lane = threadIdx.x % L;
    double part[3] = {0.0, 0.0, 0.0};
    for (long long l = (long long)sg.lb * NPB + threadIdx.x / L; l < sg.ns; l += (long long)sg.cnt * NPB) {
        const long long i = sg.n0 + l;
        T ax[kMaxNf] = {T(0), T(0), T(0), T(0)};
        if (!V.x_zero) node_spmv<L>(V.A, V.x, i, lane, ax);
        if (lane < V.nf) {
            const long long row = i + (long long)lane * V.nodes;
            const T bi = V.b[row];
            const T r = bi - pick(ax, lane);
            V.r[row] = r;
            V.rh[row] = r;
            V.p[row] = T(0);
            V.v[row] = T(0);
            if (V.x_zero) V.x[row] = T(0);
            part[0] += (double)r * (double)r;
            part[1] += (double)r * (double)r;
            part[2] += (double)bi * (double)bi;
        }
    }
    constexpr int slots[3] = {P_RHR, P_RR, P_BB};
    write_partials<3>(part, V.partials, task.block, slots);
}

// One workgroup per listed case: finalises a (re)start and writes slot `par` as the
// "previous" slot of the next iteration (which runs with parity par ^ 1).
template <typename T>
__global__ __launch_bounds__(kBlock) void k_kry_init_fin(const KrylovView<T>* __restrict__ views,
                                                         const BlockTask* __restrict__ ids, int count, int par) {
    for (int c = blockIdx.x; c < count; c += gridDim.x) {
        const KrylovView<T>& V = views[ids[c].case_id];
        if (!V.sc->req_init) continue;
        constexpr int slots[3] = {P_RHR, P_RR, P_BB};
        double tot[3];
        reduce_partials<3>(V.partials, V.nblk, slots, tot);
        if (threadIdx.x == 0) {
            KScal& S = *V.sc;
            S.bb = tot[2];
            S.tol2 = S.rtol * S.rtol * tot[2];
            if (S.atol2 > 0.0 && S.atol2 < S.tol2) S.tol2 = S.atol2;
            S.rr = tot[1];
            S.slot[par].rho = tot[0];
            S.slot[par].alpha = 1.0;
            S.slot[par].omega = 1.0;
            S.true_rr = 1;
            S.req_init = 0;
            if (tot[2] == 0.0 || tot[1] <= S.tol2) S.done = 1;
            else if (!isfinite(tot[1])) S.done = 2;
            else if (S.iters >= S.maxit) S.done = 3;
            else S.done = 0;
        }
        __syncthreads();
    }
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_kry_p(const KrylovView<T>* __restrict__ views,
                                                  const BlockTask* __restrict__ tasks, int par) {
    const BlockTask task = tasks[blockIdx.x];
    if (task.case_id < 0) return;
    const KrylovView<T>& V = views[task.case_id];
    KScal& S = *V.sc;
    if (S.done) return;
    constexpr int slots[2] = {P_RHR, P_RR};
    double tot[2];
    reduce_partials<2>(V.partials, V.nblk, slots, tot);
    const KSlot prev = S.slot[par ^ 1];
    const double beta = (tot[0] / prev.rho) * (prev.alpha / prev.omega);
    int stop = 0;
    if (tot[1] <= S.tol2) stop = 1;
    else if (!isfinite(tot[1]) || tot[0] == 0.0 || !isfinite(beta)) stop = 2;
    else if (S.iters >= S.maxit) stop = 3;
    if (stop) {
        if (task.block == 0 && threadIdx.x == 0) {
            S.rr = tot[1];
            S.done = stop;
        }
        return;
    }
    const T b = (T)beta, om = (T)prev.omega;
    const Seg sg = xcd_segment(task, V.nodes);
    const int nf = V.nf;
    for (long long l = (long long)sg.lb * kBlock + threadIdx.x; l < sg.ns; l += (long long)sg.cnt * kBlock) {
        const long long i = sg.n0 + l;
        T pv[kMaxNf], ph[kMaxNf];
#pragma unroll
        for (int f = 0; f < kMaxNf; ++f) {
            pv[f] = T(0);
            if (f < nf) {
                const long long q = (long long)f * V.nodes + i;
                pv[f] = V.r[q] + b * (V.p[q] - om * V.v[q]);
                V.p[q] = pv[f];
            }
        }
        if (V.amg) {  // ph = M^-1 p comes from the V-cycle; its pre-smoothing starts here
            if (V.mg_w0) smooth_w0<T>(V, i, pv);
            continue;
        }
        block_apply<T>(V.binv, V.nodes, nf, i, pv, ph);
#pragma unroll
        for (int f = 0; f < kMaxNf; ++f)
            if (f < nf) V.ph[(long long)f * V.nodes + i] = ph[f];
    }
    if (task.block == 0 && threadIdx.x == 0) {
        S.slot[par].rho = tot[0];
        S.rr = tot[1];
    }
}

// v = A ph over the workgroup's nodes; accumulates (rh, v).
template <int L, typename MT, typename T>
__device__ __forceinline__ void kry_v_nodes(const KrylovView<T>& V, const NodeMat<MT>& A, const Seg& sg, int lane,
                                            double (&part)[1]) {
    constexpr int NPB = kBlock / L;
    for (long long l = (long long)sg.lb * NPB + threadIdx.x / L; l < sg.ns; l += (long long)sg.cnt * NPB) {
        const long long i = sg.n0 + l;
        T y[kMaxNf];
        node_spmv<L>(A, V.ph, i, lane, y);
        if (lane < V.nf) {
            const long long row = i + (long long)lane * V.nodes;
            const T yv = pick(y, lane);
            V.v[row] = yv;
            part[0] += (double)V.rh[row] * (double)yv;
        }
    }
}

// The ELL form of an SpMV pass over a case: each wavefront of the workgroups on this XCD
// takes whole 64-node slices of the XCD's slice range; per node, emit(i, y) sees the node's
// NF row values.
template <int NF, typename T, typename F>
__device__ __forceinline__ void ell_pass(const EllMatT<T>& E, const T* __restrict__ x, const BlockTask& task,
                                         F&& emit) {
    const Seg sg = xcd_segment(task, E.nsl);
    const int lane = threadIdx.x & (kEllSlice - 1);
    for (long long l = (long long)sg.lb * kWaves + (threadIdx.x >> 6); l < sg.ns; l += (long long)sg.cnt * kWaves) {
        const long long c = sg.n0 + l;
        T y[kMaxNf];
        ell_rows<NF>(E, x, c, lane, y);
        const long long i = c * kEllSlice + lane;
        if (i < E.N) emit(i, y);
    }
}

template <typename T, int L>
__global__ __launch_bounds__(kBlock) void k_kry_v(const KrylovView<T>* __restrict__ views,
                                                  const BlockTask* __restrict__ tasks) {
    const BlockTask task = tasks[blockIdx.x];
    if (task.case_id < 0) return;
    const KrylovView<T>& V = views[task.case_id];
    if (V.sc->done) return;
    double part[1] = {0.0};
    bool done = false;
    if (V.E.vals) {
        auto emit = [&](long long i, const T (&y)[kMaxNf]) {
#pragma unroll
            for (int s = 0; s < kMaxNf; ++s)
                if (s < V.nf) {
                    const long long row = i + (long long)s * V.nodes;
                    V.v[row] = y[s];
                    part[0] += (double)V.rh[row] * (double)y[s];
                }
        };
        if (V.nf == 4) ell_pass<4>(V.E, (const T*)V.ph, task, emit);
        else ell_pass<3>(V.E, (const T*)V.ph, task, emit);
        done = true;
    }
    if (!done) {
        const Seg sg = xcd_segment(task, V.nodes);
        const int lane = threadIdx.x % L;
        kry_v_nodes<L>(V, V.A, sg, lane, part);
    }
    constexpr int slots[1] = {P_RV};
    write_partials<1>(part, V.partials, task.block, slots);
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_kry_s(const KrylovView<T>* __restrict__ views,
                                                  const BlockTask* __restrict__ tasks, int par) {
    const BlockTask task = tasks[blockIdx.x];
    if (task.case_id < 0) return;
    const KrylovView<T>& V = views[task.case_id];
    KScal& S = *V.sc;
    if (S.done) return;
    constexpr int rs[1] = {P_RV};
    double rv[1];
    reduce_partials<1>(V.partials, V.nblk, rs, rv);
    const double alpha = S.slot[par].rho / rv[0];
    if (rv[0] == 0.0 || !isfinite(alpha)) {
        if (task.block == 0 && threadIdx.x == 0) S.done = 2;
        return;
    }
    const T a = (T)alpha;
    const Seg sg = xcd_segment(task, V.nodes);
    const int nf = V.nf;
    double part[1] = {0.0};
    for (long long l = (long long)sg.lb * kBlock + threadIdx.x; l < sg.ns; l += (long long)sg.cnt * kBlock) {
        const long long i = sg.n0 + l;
        T sv[kMaxNf], sh[kMaxNf];
#pragma unroll
        for (int f = 0; f < kMaxNf; ++f) {
            sv[f] = T(0);
            if (f < nf) {
                const long long q = (long long)f * V.nodes + i;
                sv[f] = V.r[q] - a * V.v[q];
                V.s[q] = sv[f];
                part[0] += (double)sv[f] * (double)sv[f];
            }
        }
        if (V.amg) {  // sh = M^-1 s comes from the V-cycle; its pre-smoothing starts here
            if (V.mg_w0) smooth_w0<T>(V, i, sv);
            continue;
        }
        block_apply<T>(V.binv, V.nodes, nf, i, sv, sh);
#pragma unroll
        for (int f = 0; f < kMaxNf; ++f)
            if (f < nf) V.sh[(long long)f * V.nodes + i] = sh[f];
    }
    constexpr int slots[1] = {P_SS};
    write_partials<1>(part, V.partials, task.block, slots);
    if (task.block == 0 && threadIdx.x == 0) S.slot[par].alpha = alpha;
}

// t = A sh over the workgroup's nodes; accumulates (t, s) and (t, t).
template <int L, typename MT, typename T>
__device__ __forceinline__ void kry_t_nodes(const KrylovView<T>& V, const NodeMat<MT>& A, const Seg& sg, int lane,
                                            double (&part)[2]) {
    constexpr int NPB = kBlock / L;
    for (long long l = (long long)sg.lb * NPB + threadIdx.x / L; l < sg.ns; l += (long long)sg.cnt * NPB) {
        const long long i = sg.n0 + l;
        T y[kMaxNf];
        node_spmv<L>(A, V.sh, i, lane, y);
        if (lane < V.nf) {
            const long long row = i + (long long)lane * V.nodes;
            const T yv = pick(y, lane);
            V.t[row] = yv;
            part[0] += (double)yv * (double)V.s[row];
            part[1] += (double)yv * (double)yv;
        }
    }
}

template <typename T, int L>
__global__ __launch_bounds__(kBlock) void k_kry_t(const KrylovView<T>* __restrict__ views,
                                                  const BlockTask* __restrict__ tasks) {
    const BlockTask task = tasks[blockIdx.x];
    if (task.case_id < 0) return;
    const KrylovView<T>& V = views[task.case_id];
    if (V.sc->done) return;
    constexpr int ss_slot[1] = {P_SS};
    double ss[1];
    reduce_partials<1>(V.partials, V.nblk, ss_slot, ss);
    if (ss[0] <= V.sc->tol2) return;  // ||s|| converged: K_x finishes the iteration
    double part[2] = {0.0, 0.0};
    bool done = false;
    if (V.E.vals) {
        auto emit = [&](long long i, const T (&y)[kMaxNf]) {
#pragma unroll
            for (int s = 0; s < kMaxNf; ++s)
                if (s < V.nf) {
                    const long long row = i + (long long)s * V.nodes;
                    V.t[row] = y[s];
                    part[0] += (double)y[s] * (double)V.s[row];
                    part[1] += (double)y[s] * (double)y[s];
                }
        };
        if (V.nf == 4) ell_pass<4>(V.E, (const T*)V.sh, task, emit);
        else ell_pass<3>(V.E, (const T*)V.sh, task, emit);
        done = true;
    }
    if (!done) {
        const Seg sg = xcd_segment(task, V.nodes);
        const int lane = threadIdx.x % L;
        kry_t_nodes<L>(V, V.A, sg, lane, part);
    }
    constexpr int slots[2] = {P_TS, P_TT};
    write_partials<2>(part, V.partials, task.block, slots);
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_kry_x(const KrylovView<T>* __restrict__ views,
                                                  const BlockTask* __restrict__ tasks, int par) {
    const BlockTask task = tasks[blockIdx.x];
    if (task.case_id < 0) return;
    const KrylovView<T>& V = views[task.case_id];
    KScal& S = *V.sc;
    if (S.done) return;
    constexpr int sl[3] = {P_SS, P_TS, P_TT};
    double tot[3];
    reduce_partials<3>(V.partials, V.nblk, sl, tot);
    const bool sconv = tot[0] <= S.tol2;
    const double alpha = S.slot[par].alpha;
    const double omega = sconv ? 1.0 : (tot[2] == 0.0 ? 0.0 : tot[1] / tot[2]);
    const T a = (T)alpha, om = (T)omega;
    const Seg sg = xcd_segment(task, V.nodes);
    const long long nloc = (long long)V.nf * sg.ns;
    double part[2] = {0.0, 0.0};
    for (long long l = (long long)sg.lb * kBlock + threadIdx.x; l < nloc; l += (long long)sg.cnt * kBlock) {
        const long long i = seg_row(sg, V.nodes, l);
        T r;
        if (sconv) {
            V.x[i] += a * V.ph[i];
            r = V.s[i];
        } else {
            V.x[i] += a * V.ph[i] + om * V.sh[i];
            r = V.s[i] - om * V.t[i];
        }
        V.r[i] = r;
        part[0] += (double)V.rh[i] * (double)r;
        part[1] += (double)r * (double)r;
    }
    constexpr int slots[2] = {P_RHR, P_RR};
    write_partials<2>(part, V.partials, task.block, slots);
    if (task.block == 0 && threadIdx.x == 0) {
        S.slot[par].omega = omega;
        S.iters += 1;
        S.true_rr = 0;
    }
}

// Mixed precision: r = b - A x in FP64, r32 = (float) r, ||r||^2 and ||b||^2.
template <int L>
__global__ __launch_bounds__(kBlock) void k_ir_residual(const RefineView* __restrict__ views,
                                                        const BlockTask* __restrict__ tasks) {
    constexpr int NPB = kBlock / L;
    const BlockTask task = tasks[blockIdx.x];
    if (task.case_id < 0) return;
    const RefineView& V = views[task.case_id];
    const int lane = threadIdx.x % L;
    double part[2] = {0.0, 0.0};
    for (long long i = (long long)task.block * NPB + threadIdx.x / L; i < V.A.N; i += (long long)task.nblocks * NPB) {
        double ax[kMaxNf];
        node_spmv<L>(V.A, V.x, i, lane, ax);
        if (lane < V.A.nf) {
            const long long row = i + (long long)lane * V.A.N;
            const double bi = V.b[row];
            const double r = bi - pick(ax, lane);
            V.r[row] = r;
            if (V.r32) V.r32[row] = (float)r;
            part[0] += r * r;
            part[1] += bi * bi;
        }
    }
    double tot[2];
    if (reduce_last_block<2>(part, V.partials, V.counters, task.block, task.nblocks, tot) && threadIdx.x == 0) {
        V.out[0] = tot[0];
        V.out[1] = tot[1];
    }
}

__global__ __launch_bounds__(kBlock) void k_ir_update(const RefineView* __restrict__ views,
                                                      const BlockTask* __restrict__ tasks) {
    const BlockTask task = tasks[blockIdx.x];
    if (task.case_id < 0) return;
    const RefineView& V = views[task.case_id];
    for (long long i = (long long)task.block * kBlock + threadIdx.x; i < V.n; i += (long long)task.nblocks * kBlock)
        V.x[i] += (double)V.d32[i];
}

// Launchers ---------------------------------------------------------------------------

// lanes: the lanes per node of the node-block SpMV (4, 8 or 16)
#define PFEM_DISPATCH_L(lanes, CALL)                      \
    switch (lanes) {                                       \
        case 4: { constexpr int LL = 4; CALL; } break;     \
        case 8: { constexpr int LL = 8; CALL; } break;     \
        default: { constexpr int LL = 16; CALL; } break;   \
    }

template <typename T>
void launch_kry_init(int lanes, const KrylovView<T>* v, const BlockTask* t, int nt, const BlockTask* ids, int nids,
                     int par, hipStream_t s) {
    PFEM_DISPATCH_L(lanes, (k_kry_init<T, LL><<<nt, kBlock, 0, s>>>(v, t)));
    k_kry_init_fin<T><<<nids < 1024 ? nids : 1024, kBlock, 0, s>>>(v, ids, nids, par);
}


// One kernel of an iteration (phase 0 K_p, 1 K_v, 2 K_s, 3 K_t, 4 K_x) with parity par, so
// that a preconditioner's launches can sit between them; ev (optional) brackets K_v.
template <typename T>
void launch_kry_phase(int phase, int lanes, const KrylovView<T>* v, const BlockTask* t, int nt, int par,
                      hipStream_t s, hipEvent_t* ev) {
    switch (phase) {
        case 0: k_kry_p<T><<<nt, kBlock, 0, s>>>(v, t, par); break;
        case 1:
            // timed: the events ride on the dispatch itself (start and end of the kernel,
            // as a profiler sees them), not as separate packets around it
            if (ev) {
                PFEM_DISPATCH_L(lanes, (hipExtLaunchKernelGGL(k_kry_v<T, LL>, dim3(nt), dim3(kBlock), 0, s, ev[0],
                                                              ev[1], 0, v, t)));
            } else {
                PFEM_DISPATCH_L(lanes, (k_kry_v<T, LL><<<nt, kBlock, 0, s>>>(v, t)));
            }
            break;
        case 2: k_kry_s<T><<<nt, kBlock, 0, s>>>(v, t, par); break;
        case 3: PFEM_DISPATCH_L(lanes, (k_kry_t<T, LL><<<nt, kBlock, 0, s>>>(v, t))); break;
        default: k_kry_x<T><<<nt, kBlock, 0, s>>>(v, t, par); break;
    }
}

// y = A x over the sliced ELL copy (one wavefront per 64-node slice), the SpMV the BiCGSTAB
// kernels run; for checking the copy against the node-block values.
template <int NF>
__global__ __launch_bounds__(kBlock) void k_spmv_ell(const EllMatT<double> E, const double* __restrict__ x,
                                                     double* __restrict__ y) {
    const int lane = threadIdx.x & (kEllSlice - 1);
    for (long long c = (long long)blockIdx.x * kWaves + (threadIdx.x >> 6); c < E.nsl; c += (long long)gridDim.x * kWaves) {
        double r[kMaxNf];
        ell_rows<NF, double, double>(E, x, c, lane, r);
        const long long i = c * kEllSlice + lane;
        if (i < E.N)
#pragma unroll
            for (int s = 0; s < NF; ++s) y[i + (long long)s * E.N] = r[s];
    }
}

void launch_spmv_ell(const EllMatT<double>& E, const double* x, double* y, int nblocks, hipStream_t s) {
    if (E.nf == 3) k_spmv_ell<3><<<nblocks, kBlock, 0, s>>>(E, x, y);
    else k_spmv_ell<4><<<nblocks, kBlock, 0, s>>>(E, x, y);
}

template <typename T>
void launch_spmv(int lanes, const NodeMat<T>& A, const T* x, T* y, int nblocks, hipStream_t s) {
    PFEM_DISPATCH_L(lanes, (k_spmv<T, LL><<<nblocks, kBlock, 0, s>>>(A, x, y)));
}

void launch_ir_residual(int lanes, const RefineView* v, const BlockTask* t, int nt, hipStream_t s) {
    PFEM_DISPATCH_L(lanes, (k_ir_residual<LL><<<nt, kBlock, 0, s>>>(v, t)));
}

void launch_ir_update(const RefineView* v, const BlockTask* t, int nt, hipStream_t s) {
    k_ir_update<<<nt, kBlock, 0, s>>>(v, t);
}

template void launch_kry_init<double>(int, const KrylovView<double>*, const BlockTask*, int, const BlockTask*, int,
                                      int, hipStream_t);
template void launch_kry_init<float>(int, const KrylovView<float>*, const BlockTask*, int, const BlockTask*, int, int,
                                     hipStream_t);
template void launch_kry_phase<double>(int, int, const KrylovView<double>*, const BlockTask*, int, int, hipStream_t,
                                       hipEvent_t*);
template void launch_kry_phase<float>(int, int, const KrylovView<float>*, const BlockTask*, int, int, hipStream_t,
                                      hipEvent_t*);
template void launch_spmv<double>(int, const NodeMat<double>&, const double*, double*, int, hipStream_t);
template void launch_spmv<float>(int, const NodeMat<float>&, const float*, float*, int, hipStream_t);

}  // namespace pfem
