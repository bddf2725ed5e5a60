// K3 (CSR SpMV) and K5 (Jacobi-preconditioned BiCGSTAB) for a batch of cases, plus the
// FP64 iterative-refinement kernels of the mixed-precision solve.
//
// Replaces SparseLU<SparseMatrix<double>, COLAMDOrdering<int>> (analyzePattern / factorize /
// solve at MomContEquationPSPG.inl:281-290) with right-preconditioned BiCGSTAB
// (van der Vorst 1992).  One iteration is five launches:
//   p  = r + beta (p - omega v);  ph = D^-1 p
//   v  = A ph                      -> (rh, v)        -> alpha = rho / (rh, v)
//   s  = r - alpha v;  sh = D^-1 s -> (s, s)         -> early exit when ||s|| <= tol
//   t  = A sh                      -> (t, s), (t, t) -> omega
//   x += alpha ph + omega sh;  r = s - omega t -> (rh, r), (r, r) -> beta, rho, convergence
// Every dot product is accumulated in FP64 and reduced deterministically; the scalars
// are finalised on the device by the last workgroup of each reduction, so the host
// only reads the convergence flag every few iterations.
//
// SpMV: L lanes per row (L | 64), rows grid-strided over the case's workgroups; the
// L lanes of a row read consecutive values and column indices (coalesced) and reduce
// with cross-lane shuffles.
#include "common.hpp"
#include "views.hpp"

namespace pfem {

template <typename T, int L>
__device__ __forceinline__ T row_dot(const int* __restrict__ rp, const int* __restrict__ ci, const T* __restrict__ va,
                                     const T* __restrict__ x, long long row, int lane) {
    const int beg = rp[row], end = rp[row + 1];
    T acc = 0;
    for (int k = beg + lane; k < end; k += L) acc += va[k] * x[ci[k]];
#pragma unroll
    for (int o = L / 2; o > 0; o >>= 1) acc += __shfl_xor(acc, o, L);
    return acc;
}

// y = A x (plain SpMV; used for the roofline measurement and the tests).
template <typename T, int L>
__global__ __launch_bounds__(kBlock) void k_spmv(int n, const int* __restrict__ rp, const int* __restrict__ ci,
                                                 const T* __restrict__ va, const T* __restrict__ x, T* __restrict__ y,
                                                 int nblocks) {
    constexpr int RPB = kBlock / L;
    const int lane = threadIdx.x % L;
    for (long long row = (long long)blockIdx.x * RPB + threadIdx.x / L; row < n; row += (long long)nblocks * RPB) {
        const T acc = row_dot<T, L>(rp, ci, va, x, row, lane);
        if (lane == 0) y[row] = acc;
    }
}

template <typename T, int L>
__global__ __launch_bounds__(kBlock) void k_kry_init(const KrylovView<T>* __restrict__ views,
                                                     const BlockTask* __restrict__ tasks) {
    constexpr int RPB = kBlock / L;
    const BlockTask task = tasks[blockIdx.x];
    if (task.case_id < 0) return;
    const KrylovView<T>& V = views[task.case_id];
    if (!V.sc->req_init) return;  // only cases the host asked to (re)start
    const int lane = threadIdx.x % L;
    double part[2] = {0.0, 0.0};
    for (long long row = (long long)task.block * RPB + threadIdx.x / L; row < V.n; row += (long long)task.nblocks * RPB) {
        const T ax = V.x_zero ? T(0) : row_dot<T, L>(V.row_ptr, V.col_idx, V.vals, V.x, row, lane);
        if (lane == 0) {
            const T bi = V.b[row];
            const T r = bi - ax;
            V.r[row] = r;
            V.rh[row] = r;
            V.p[row] = T(0);
            V.v[row] = T(0);
            if (V.x_zero) V.x[row] = T(0);
            part[0] += (double)r * (double)r;
            part[1] += (double)bi * (double)bi;
        }
    }
    double tot[2];
    if (reduce_last_block<2>(part, V.partials, V.counters, task.block, task.nblocks, tot) && threadIdx.x == 0) {
        KScal& S = *V.sc;
        S.rr = tot[0];
        S.bb = tot[1];
        S.tol2 = S.rtol * S.rtol * tot[1];
        S.rho = tot[0];
        S.alpha = 1.0;
        S.omega = 1.0;
        S.beta = 0.0;
        S.s_conv = 0;
        S.req_init = 0;
        if (tot[1] == 0.0 || tot[0] <= S.tol2) S.done = 1;
        else if (!isfinite(tot[0])) S.done = 2;
        else if (S.iters >= S.maxit) S.done = 3;
        else S.done = 0;
    }
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_kry_p(const KrylovView<T>* __restrict__ views,
                                                  const BlockTask* __restrict__ tasks) {
    const BlockTask task = tasks[blockIdx.x];
    if (task.case_id < 0) return;
    const KrylovView<T>& V = views[task.case_id];
    const KScal& S = *V.sc;
    if (S.done) return;
    const T beta = (T)S.beta, omega = (T)S.omega;
    for (long long i = (long long)task.block * kBlock + threadIdx.x; i < V.n; i += (long long)task.nblocks * kBlock) {
        const T p = V.r[i] + beta * (V.p[i] - omega * V.v[i]);
        V.p[i] = p;
        V.ph[i] = V.dinv[i] * p;
    }
}

template <typename T, int L>
__global__ __launch_bounds__(kBlock) void k_kry_spmv_v(const KrylovView<T>* __restrict__ views,
                                                       const BlockTask* __restrict__ tasks) {
    constexpr int RPB = kBlock / L;
    const BlockTask task = tasks[blockIdx.x];
    if (task.case_id < 0) return;
    const KrylovView<T>& V = views[task.case_id];
    if (V.sc->done) return;
    const int lane = threadIdx.x % L;
    double part[1] = {0.0};
    for (long long row = (long long)task.block * RPB + threadIdx.x / L; row < V.n; row += (long long)task.nblocks * RPB) {
        const T y = row_dot<T, L>(V.row_ptr, V.col_idx, V.vals, V.ph, row, lane);
        if (lane == 0) {
            V.v[row] = y;
            part[0] += (double)V.rh[row] * (double)y;
        }
    }
    double tot[1];
    if (reduce_last_block<1>(part, V.partials, V.counters, task.block, task.nblocks, tot) && threadIdx.x == 0) {
        KScal& S = *V.sc;
        S.rv = tot[0];
        const double alpha = S.rho / tot[0];
        if (tot[0] == 0.0 || !isfinite(alpha)) S.done = 2;
        else S.alpha = alpha;
    }
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_kry_s(const KrylovView<T>* __restrict__ views,
                                                  const BlockTask* __restrict__ tasks) {
    const BlockTask task = tasks[blockIdx.x];
    if (task.case_id < 0) return;
    const KrylovView<T>& V = views[task.case_id];
    if (V.sc->done) return;
    const T alpha = (T)V.sc->alpha;
    double part[1] = {0.0};
    for (long long i = (long long)task.block * kBlock + threadIdx.x; i < V.n; i += (long long)task.nblocks * kBlock) {
        const T s = V.r[i] - alpha * V.v[i];
        V.s[i] = s;
        V.sh[i] = V.dinv[i] * s;
        part[0] += (double)s * (double)s;
    }
    double tot[1];
    if (reduce_last_block<1>(part, V.partials, V.counters, task.block, task.nblocks, tot) && threadIdx.x == 0) {
        KScal& S = *V.sc;
        S.ss = tot[0];
        S.s_conv = (tot[0] <= S.tol2) ? 1 : 0;
    }
}

template <typename T, int L>
__global__ __launch_bounds__(kBlock) void k_kry_spmv_t(const KrylovView<T>* __restrict__ views,
                                                       const BlockTask* __restrict__ tasks) {
    constexpr int RPB = kBlock / L;
    const BlockTask task = tasks[blockIdx.x];
    if (task.case_id < 0) return;
    const KrylovView<T>& V = views[task.case_id];
    if (V.sc->done || V.sc->s_conv) return;
    const int lane = threadIdx.x % L;
    double part[2] = {0.0, 0.0};
    for (long long row = (long long)task.block * RPB + threadIdx.x / L; row < V.n; row += (long long)task.nblocks * RPB) {
        const T y = row_dot<T, L>(V.row_ptr, V.col_idx, V.vals, V.sh, row, lane);
        if (lane == 0) {
            V.t[row] = y;
            part[0] += (double)y * (double)V.s[row];
            part[1] += (double)y * (double)y;
        }
    }
    double tot[2];
    if (reduce_last_block<2>(part, V.partials, V.counters, task.block, task.nblocks, tot) && threadIdx.x == 0) {
        KScal& S = *V.sc;
        S.omega = (tot[1] == 0.0) ? 0.0 : tot[0] / tot[1];
    }
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_kry_x(const KrylovView<T>* __restrict__ views,
                                                  const BlockTask* __restrict__ tasks) {
    const BlockTask task = tasks[blockIdx.x];
    if (task.case_id < 0) return;
    const KrylovView<T>& V = views[task.case_id];
    if (V.sc->done) return;
    const bool sconv = V.sc->s_conv != 0;
    const T alpha = (T)V.sc->alpha, omega = (T)V.sc->omega;
    double part[2] = {0.0, 0.0};
    for (long long i = (long long)task.block * kBlock + threadIdx.x; i < V.n; i += (long long)task.nblocks * kBlock) {
        if (sconv) {
            V.x[i] += alpha * V.ph[i];
        } else {
            V.x[i] += alpha * V.ph[i] + omega * V.sh[i];
            const T r = V.s[i] - omega * V.t[i];
            V.r[i] = r;
            part[0] += (double)V.rh[i] * (double)r;
            part[1] += (double)r * (double)r;
        }
    }
    double tot[2];
    if (reduce_last_block<2>(part, V.partials, V.counters, task.block, task.nblocks, tot) && threadIdx.x == 0) {
        KScal& S = *V.sc;
        S.iters += 1;
        if (sconv) {
            S.rr = S.ss;
            S.done = 1;
        } else {
            S.rr = tot[1];
            if (tot[1] <= S.tol2) {
                S.done = 1;
            } else if (!isfinite(tot[1]) || !isfinite(tot[0])) {
                S.done = 2;
            } else if (S.omega == 0.0 || tot[0] == 0.0) {
                S.done = 2;  // breakdown: the host restarts from the current x
            } else {
                S.beta = (tot[0] / S.rho) * (S.alpha / S.omega);
                S.rho = tot[0];
                if (S.iters >= S.maxit) S.done = 3;
            }
        }
    }
}

// Mixed precision: r = b - A x in FP64, r32 = (float) r, ||r||^2 and ||b||^2.
template <int L>
__global__ __launch_bounds__(kBlock) void k_ir_residual(const RefineView* __restrict__ views,
                                                        const BlockTask* __restrict__ tasks) {
    constexpr int RPB = kBlock / L;
    const BlockTask task = tasks[blockIdx.x];
    if (task.case_id < 0) return;
    const RefineView& V = views[task.case_id];
    const int lane = threadIdx.x % L;
    double part[2] = {0.0, 0.0};
    for (long long row = (long long)task.block * RPB + threadIdx.x / L; row < V.n; row += (long long)task.nblocks * RPB) {
        const double ax = row_dot<double, L>(V.row_ptr, V.col_idx, V.vals, V.x, row, lane);
        if (lane == 0) {
            const double bi = V.b[row];
            const double r = bi - ax;
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
// ev (optional): a pair of events recorded around the first SpMV of the iteration, for
// live timing of that kernel.
template <typename T, int L>
static void iter_launch(const KrylovView<T>* v, const BlockTask* t, int nt, hipStream_t s, hipEvent_t* ev) {
    k_kry_p<T><<<nt, kBlock, 0, s>>>(v, t);
    if (ev) (void)hipEventRecord(ev[0], s);
    k_kry_spmv_v<T, L><<<nt, kBlock, 0, s>>>(v, t);
    if (ev) (void)hipEventRecord(ev[1], s);
    k_kry_s<T><<<nt, kBlock, 0, s>>>(v, t);
    k_kry_spmv_t<T, L><<<nt, kBlock, 0, s>>>(v, t);
    k_kry_x<T><<<nt, kBlock, 0, s>>>(v, t);
}

#define PFEM_DISPATCH_L(lanes, CALL)                      \
    switch (lanes) {                                       \
        case 4: { constexpr int LL = 4; CALL; } break;     \
        case 8: { constexpr int LL = 8; CALL; } break;     \
        case 16: { constexpr int LL = 16; CALL; } break;   \
        case 32: { constexpr int LL = 32; CALL; } break;   \
        default: { constexpr int LL = 64; CALL; } break;   \
    }

template <typename T>
void launch_kry_init(int lanes, const KrylovView<T>* v, const BlockTask* t, int nt, hipStream_t s) {
    PFEM_DISPATCH_L(lanes, (k_kry_init<T, LL><<<nt, kBlock, 0, s>>>(v, t)));
}

template <typename T>
void launch_kry_iters(int lanes, int iters, const KrylovView<T>* v, const BlockTask* t, int nt, hipStream_t s,
                      hipEvent_t* ev) {
    for (int it = 0; it < iters; ++it)
        PFEM_DISPATCH_L(lanes, (iter_launch<T, LL>(v, t, nt, s, ev ? ev + 2 * it : nullptr)));
}

template <typename T>
void launch_spmv(int lanes, int n, const int* rp, const int* ci, const T* va, const T* x, T* y, int nblocks,
                 hipStream_t s) {
    PFEM_DISPATCH_L(lanes, (k_spmv<T, LL><<<nblocks, kBlock, 0, s>>>(n, rp, ci, va, x, y, nblocks)));
}

void launch_ir_residual(int lanes, const RefineView* v, const BlockTask* t, int nt, hipStream_t s) {
    PFEM_DISPATCH_L(lanes, (k_ir_residual<LL><<<nt, kBlock, 0, s>>>(v, t)));
}

void launch_ir_update(const RefineView* v, const BlockTask* t, int nt, hipStream_t s) {
    k_ir_update<<<nt, kBlock, 0, s>>>(v, t);
}

template void launch_kry_init<double>(int, const KrylovView<double>*, const BlockTask*, int, hipStream_t);
template void launch_kry_init<float>(int, const KrylovView<float>*, const BlockTask*, int, hipStream_t);
template void launch_kry_iters<double>(int, int, const KrylovView<double>*, const BlockTask*, int, hipStream_t,
                                       hipEvent_t*);
template void launch_kry_iters<float>(int, int, const KrylovView<float>*, const BlockTask*, int, hipStream_t,
                                      hipEvent_t*);
template void launch_spmv<double>(int, int, const int*, const int*, const double*, const double*, double*, int,
                                  hipStream_t);
template void launch_spmv<float>(int, int, const int*, const int*, const float*, const float*, float*, int,
                                 hipStream_t);

}  // namespace pfem
