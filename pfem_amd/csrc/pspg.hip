// K1 (element assembly) and K2 (boundary conditions) of the PSPG momentum-continuity
// system, MomContEquationPSPG.inl:7-235, for a batch of cases.
//
// K1 runs one element per lane.  The element geometry (J, detJ, inverse J; Element.cpp:15-135)
// and the P1 shape-function gradients (MatricesBuilder.inl:93-128) stay in registers and
// every entry of Ae (MomContEquationPSPG.inl:42) is evaluated from closed forms of the
// P1 element matrices, which the reference's Gauss rules integrate exactly:
//   M_ij  = rho |e| (1 + d_ij) / ((d+1)(d+2))                       getM   (MatricesBuilder.inl:217)
//   K_(is)(jt) = mu |e| (d_st gradNi.gradNj + d_tNi d_sNj)          getK   (:277, ddev diag(2,..,1,..))
//   -D^T_(is)(j) = -|e|/(d+1) d_sNi                                 getD   (:293)
//   (C tau/dt + D)_(i)(jt) = |e|/(d+1) (tau/dt d_tNi + d_tNj)       getC   (:325)
//   tau L_ij = tau/rho |e| gradNi.gradNj                            getL   (:309)
//   F_(is) = rho |e| g_s/(d+1),  H_i = tau |e| gradNi.g             getF, getH (:341, :373)
// and scattered with FP64 atomics into the GPU-built CSR (pattern.hpp) at
// row_start(i, s) + t*deg(i) + rank_i(j).  Momentum rows of boundary nodes are skipped
// as at MomContEquationPSPG.inl:68 and receive the unit diagonal of :119-126.
#include "common.hpp"
#include "views.hpp"

namespace pfem {

template <int D>
struct Geom {
    double gN[D][D + 1];  // gradient of N_i, component k: gN[k][i]
    double vol;           // |e| = detJ * refSize
    double detJ;
};

template <int D>
__device__ __forceinline__ void element_geometry(const double* __restrict__ pos, int N, const int (&nd)[D + 1],
                                                 Geom<D>& G) {
    double x[D][D + 1];
#pragma unroll
    for (int k = 0; k < D; ++k)
#pragma unroll
        for (int a = 0; a <= D; ++a) x[k][a] = pos[(long long)k * N + nd[a]];
    double J[D][D];
#pragma unroll
    for (int r = 0; r < D; ++r)
#pragma unroll
        for (int c = 0; c < D; ++c) J[r][c] = x[r][c + 1] - x[r][0];
    if constexpr (D == 2) {
        const double dj = J[0][0] * J[1][1] - J[1][0] * J[0][1];
        const double inv[2][2] = {{J[1][1] / dj, -J[0][1] / dj}, {-J[1][0] / dj, J[0][0] / dj}};
#pragma unroll
        for (int k = 0; k < 2; ++k) {
            G.gN[k][0] = -inv[0][k] - inv[1][k];
            G.gN[k][1] = inv[0][k];
            G.gN[k][2] = inv[1][k];
        }
        G.detJ = dj;
        G.vol = dj * 0.5;
    } else {
        const double dj = J[0][0] * J[1][1] * J[2][2] + J[0][1] * J[1][2] * J[2][0] + J[0][2] * J[1][0] * J[2][1] -
                          J[2][0] * J[1][1] * J[0][2] - J[2][1] * J[1][2] * J[0][0] - J[2][2] * J[1][0] * J[0][1];
        double inv[3][3];
        inv[0][0] = (J[1][1] * J[2][2] - J[1][2] * J[2][1]) / dj;
        inv[0][1] = (J[2][1] * J[0][2] - J[2][2] * J[0][1]) / dj;
        inv[0][2] = (J[0][1] * J[1][2] - J[0][2] * J[1][1]) / dj;
        inv[1][0] = (J[2][0] * J[1][2] - J[1][0] * J[2][2]) / dj;
        inv[1][1] = (J[0][0] * J[2][2] - J[2][0] * J[0][2]) / dj;
        inv[1][2] = (J[1][0] * J[0][2] - J[0][0] * J[1][2]) / dj;
        inv[2][0] = (J[1][0] * J[2][1] - J[2][0] * J[1][1]) / dj;
        inv[2][1] = (J[2][0] * J[0][1] - J[0][0] * J[2][1]) / dj;
        inv[2][2] = (J[0][0] * J[1][1] - J[1][0] * J[0][1]) / dj;
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            G.gN[k][0] = -inv[0][k] - inv[1][k] - inv[2][k];
            G.gN[k][1] = inv[0][k];
            G.gN[k][2] = inv[1][k];
            G.gN[k][3] = inv[2][k];
        }
        G.detJ = dj;
        G.vol = dj * 0.16666666666666666666666666666667;
    }
}

__device__ __forceinline__ uint32_t rank_of(const uint4 r, int q) {
    const uint32_t w = (q < 4) ? r.x : (q < 8) ? r.y : (q < 12) ? r.z : r.w;
    return (w >> (8 * (q & 3))) & 0xffu;
}

// One element per lane; accumulates A (FP64 atomics into the CSR) and b.
template <int D>
__global__ __launch_bounds__(kBlock) void k_pspg_assemble(const PspgView* __restrict__ views,
                                                          const BlockTask* __restrict__ tasks) {
    constexpr int NE = D + 1;
    const BlockTask task = tasks[blockIdx.x];
    if (task.case_id < 0) return;
    const PspgView& V = views[task.case_id];
    const int N = V.N, E = V.E;
    const double inv_ne = 1.0 / NE;
    const double mcoef = 1.0 / (double)((D + 1) * (D + 2));
    for (long long e = (long long)task.block * kBlock + threadIdx.x; e < E; e += (long long)task.nblocks * kBlock) {
        int nd[NE];
#pragma unroll
        for (int a = 0; a < NE; ++a) nd[a] = V.conn[(long long)a * E + e];
        Geom<D> G;
        element_geometry<D>(V.pos, N, nd, G);
        // tau: MomContEquationPSPG.inl:238-259 (node velocities of the current iterate)
        const double h = sqrt((D == 2 ? 0.5 : 0.16666666666666666666666666666667) * G.detJ / M_PI);
        double U = 0;
        double vp[D][NE];  // previous-step velocities (be)
#pragma unroll
        for (int a = 0; a < NE; ++a) {
            double u2 = 0;
#pragma unroll
            for (int k = 0; k < D; ++k) {
                const double vi = V.q_iter[(long long)k * N + nd[a]];
                u2 += vi * vi;
                vp[k][a] = V.q_prev[(long long)k * N + nd[a]];
            }
            U += sqrt(u2);
        }
        U /= (D + 1);
        const double dt = V.dt;
        const double visc = 4 * V.mu / (h * h * V.rho);
        const double tau = 1 / sqrt((2 / dt) * (2 / dt) + (2 * U / h) * (2 * U / h) + 9 * visc * visc);
        const double vol = G.vol;
        const double mu_vol = V.mu * vol;
        const double vol_ne = vol * inv_ne;
        const double tau_dt = tau / dt;
        const uint4 rk = V.ranks[e];
        double vsum[D];
#pragma unroll
        for (int k = 0; k < D; ++k) {
            double s = 0;
#pragma unroll
            for (int b = 0; b < NE; ++b) s += vp[k][b];
            vsum[k] = s;
        }
#pragma unroll
        for (int a = 0; a < NE; ++a) {
            const int4 inf = V.node_info[nd[a]];
            const int deg = inf.z;
            const bool mom = (inf.w & (NODE_BOUND | NODE_FREE)) == 0;
            const long long cont_base = (long long)D * V.mom_nnz + inf.y;
#pragma unroll
            for (int b = 0; b < NE; ++b) {
                const uint32_t r = rank_of(rk, a * NE + b);
                double gdot = 0;
#pragma unroll
                for (int k = 0; k < D; ++k) gdot += G.gN[k][a] * G.gN[k][b];
                if (mom) {
                    const double mab = V.rho * vol * ((a == b) ? 2.0 : 1.0) * mcoef / dt;
#pragma unroll
                    for (int s = 0; s < D; ++s) {
                        double* row = V.vals + (long long)s * V.mom_nnz + inf.x + r;
#pragma unroll
                        for (int t = 0; t < D; ++t) {
                            double val = mu_vol * G.gN[t][a] * G.gN[s][b];
                            if (s == t) val += mab + mu_vol * gdot;
                            atomicAdd(row + (long long)t * deg, val);
                        }
                        atomicAdd(row + (long long)D * deg, -vol_ne * G.gN[s][a]);
                    }
                }
                double* crow = V.vals + cont_base + r;
#pragma unroll
                for (int t = 0; t < D; ++t)
                    atomicAdd(crow + (long long)t * deg, vol_ne * (tau_dt * G.gN[t][a] + G.gN[t][b]));
                atomicAdd(crow + (long long)D * deg, tau / V.rho * vol * gdot);
            }
            // be (MomContEquationPSPG.inl:53): F + Me/dt vPrev ; H + C tau/dt vPrev
#pragma unroll
            for (int s = 0; s < D; ++s) {
                double mv = 0;
#pragma unroll
                for (int b = 0; b < NE; ++b) mv += ((a == b) ? 2.0 : 1.0) * vp[s][b];
                const double be = V.rho * vol_ne * V.g[s] + V.rho * vol * mcoef / dt * mv;
                atomicAdd(&V.b[(long long)s * N + nd[a]], be);
            }
            double hg = 0, cv = 0;
#pragma unroll
            for (int k = 0; k < D; ++k) {
                hg += G.gN[k][a] * V.g[k];
                cv += G.gN[k][a] * vsum[k];
            }
            atomicAdd(&V.b[(long long)D * N + nd[a]], tau * vol * hg + tau_dt * vol_ne * cv);
        }
    }
}

// Zero A and b of every case before the atomic accumulation (a fill kernel rather than
// hipMemsetAsync so that the launch is safe inside captured graphs).
__global__ __launch_bounds__(kBlock) void k_pspg_zero(const PspgView* __restrict__ views,
                                                      const BlockTask* __restrict__ tasks) {
    const BlockTask task = tasks[blockIdx.x];
    if (task.case_id < 0) return;
    const PspgView& V = views[task.case_id];
    const long long n = (long long)(V.dim + 1) * V.N;
    const long long stride = (long long)task.nblocks * kBlock;
    for (long long k = (long long)task.block * kBlock + threadIdx.x; k < V.nnz; k += stride) V.vals[k] = 0.0;
    for (long long k = (long long)task.block * kBlock + threadIdx.x; k < n; k += stride) V.b[k] = 0.0;
}

// Unit diagonals of boundary/free momentum rows and free continuity rows
// (MomContEquationPSPG.inl:108-128), then the nodal right-hand side of m_applyBCPSPG
// (MomContEquationPSPG.inl:190-218): free nodes b_p = 0 and, if not boundary,
// b_u = qPrev_u + dt g; Dirichlet nodes b_u = v_bc.
__global__ __launch_bounds__(kBlock) void k_pspg_nodes(const PspgView* __restrict__ views,
                                                       const BlockTask* __restrict__ tasks) {
    const BlockTask task = tasks[blockIdx.x];
    if (task.case_id < 0) return;
    const PspgView& V = views[task.case_id];
    const int N = V.N, D = V.dim;
    for (int i = task.block * kBlock + threadIdx.x; i < N; i += task.nblocks * kBlock) {
        const int4 inf = V.node_info[i];
        const uint8_t f = (uint8_t)inf.w;
        const bool isfree = f & NODE_FREE, bound = f & NODE_BOUND;
        if (bound || isfree)
            for (int s = 0; s < D; ++s) V.vals[(long long)s * V.mom_nnz + inf.x] = 1.0;
        if (isfree) {
            V.vals[(long long)D * V.mom_nnz + inf.y] = 1.0;
            V.b[(long long)D * N + i] = 0.0;
            if (!bound)
                for (int s = 0; s < D; ++s) V.b[(long long)s * N + i] = V.q_prev[(long long)s * N + i] + V.dt * V.g[s];
        }
        if (bound && (f & NODE_DIRICHLET))
            for (int s = 0; s < D; ++s) V.b[(long long)s * N + i] = V.dir_vals[(long long)s * N + i];
    }
}

// Dirichlet column elimination (MomContEquationPSPG.inl:216-228): b[row] -= A(row, col) v_bc;
// A(row, col) = 0, kept as an explicit zero.
__global__ __launch_bounds__(kBlock) void k_pspg_dirichlet(const PspgView* __restrict__ views,
                                                           const BlockTask* __restrict__ tasks) {
    const BlockTask task = tasks[blockIdx.x];
    if (task.case_id < 0) return;
    const PspgView& V = views[task.case_id];
    for (int k = task.block * kBlock + threadIdx.x; k < V.n_dir; k += task.nblocks * kBlock) {
        const int4 ent = V.dir[k];
        const double vbc = V.dir_vals[ent.z];
        const double a = V.vals[ent.x];
        if (vbc != 0.0) atomicAdd(&V.b[ent.y], -a * vbc);
        V.vals[ent.x] = 0.0;
    }
}

// Jacobi inverse diagonal (FP64, plus the FP32 copies of A and D^-1 when requested).
__global__ __launch_bounds__(kBlock) void k_pspg_precond(const PspgView* __restrict__ views,
                                                         const BlockTask* __restrict__ tasks) {
    const BlockTask task = tasks[blockIdx.x];
    if (task.case_id < 0) return;
    const PspgView& V = views[task.case_id];
    const long long n = (long long)(V.dim + 1) * V.N;
    const long long stride = (long long)task.nblocks * kBlock;
    for (long long r = (long long)task.block * kBlock + threadIdx.x; r < n; r += stride) {
        const double d = 1.0 / V.vals[V.diag_pos[r]];
        V.dinv[r] = d;
        if (V.dinv32) V.dinv32[r] = (float)d;
    }
    if (V.vals32)
        for (long long k = (long long)task.block * kBlock + threadIdx.x; k < V.nnz; k += stride)
            V.vals32[k] = (float)V.vals[k];
}

// Launchers ---------------------------------------------------------------------------
void launch_pspg_assemble(int dim, const PspgView* views, const BlockTask* tasks, int ntasks, hipStream_t s) {
    k_pspg_zero<<<ntasks, kBlock, 0, s>>>(views, tasks);
    if (dim == 2) k_pspg_assemble<2><<<ntasks, kBlock, 0, s>>>(views, tasks);
    else k_pspg_assemble<3><<<ntasks, kBlock, 0, s>>>(views, tasks);
}

void launch_pspg_bc(const PspgView* views, const BlockTask* tasks, int ntasks, hipStream_t s) {
    k_pspg_nodes<<<ntasks, kBlock, 0, s>>>(views, tasks);
    k_pspg_dirichlet<<<ntasks, kBlock, 0, s>>>(views, tasks);
}

void launch_pspg_precond(const PspgView* views, const BlockTask* tasks, int ntasks, hipStream_t s) {
    k_pspg_precond<<<ntasks, kBlock, 0, s>>>(views, tasks);
}

}  // namespace pfem
