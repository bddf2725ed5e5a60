// K1 (assembly) and K2 (boundary conditions) of the PSPG momentum-continuity system,
// MomContEquationPSPG.inl:7-235, for a batch of cases.
//
// K1 is a gather assembly: every stored value is computed by the lane that owns it and
// written exactly once, with no atomics.  For node-graph entry (i, j) the elements that
// contain both nodes come from the pair-element incidence built with the pattern (sorted
// by element index, so the element contributions of an entry are summed in the same
// order as Eigen's duplicate sum of the reference's triplet list).  Each element's
// geometry (J, detJ, inverse J; Element.cpp:15-135), its P1 shape-function gradients
// (MatricesBuilder.inl:93-128) and tau (MomContEquationPSPG.inl:238-259) are recomputed
// in registers for every entry it touches; the entries of Ae (MomContEquationPSPG.inl:42)
// follow from closed forms of the P1 element matrices, which the reference's Gauss rules
// integrate exactly:
//   M_ij  = rho |e| (1 + d_ij) / ((d+1)(d+2))                       getM   (MatricesBuilder.inl:217)
//   K_(is)(jt) = mu |e| (d_st gradNi.gradNj + d_tNi d_sNj)          getK   (:277, ddev diag(2,..,1,..))
//   -D^T_(is)(j) = -|e|/(d+1) d_sNi                                 getD   (:293)
//   (C tau/dt + D)_(i)(jt) = |e|/(d+1) (tau/dt d_tNi + d_tNj)       getC   (:325)
//   tau L_ij = tau/rho |e| gradNi.gradNj                            getL   (:309)
//   F_(is) = rho |e| g_s/(d+1),  H_i = tau |e| gradNi.g             getF, getH (:341, :373)
// Off-diagonal node pairs take one lane each; the diagonal block and the right-hand side
// of a node (every element around it) are spread over the node's lane group and reduced
// with cross-lane shuffles.  Momentum rows of boundary nodes are not assembled
// (MomContEquationPSPG.inl:68); K2 gives them, and every row of a free node, the unit
// diagonal of MomContEquationPSPG.inl:108-128.
//
// K2 is node-major and also atomic-free: the Dirichlet column elimination
// (MomContEquationPSPG.inl:216-228) is done row by row — the owner of row (j, s) walks its
// sorted neighbours and their velocity components, which subtracts into b[(j, s)] in the
// reference's order (Dirichlet node index, then component).
#include "common.hpp"
#include "views.hpp"

namespace pfem {

constexpr int kAsmLanes = 16;  // lanes per node in the assembly kernels

template <int D>
struct Geom {
    double gN[D][D + 1];  // gradient of N_i, component k: gN[k][i]
    double vol;           // |e| = detJ * refSize
    double tau;
};

// Geometry and tau of element e (nodes nd), MomContEquationPSPG.inl:238-259.
template <int D>
__device__ __forceinline__ void element_eval(const PspgView& V, const int (&nd)[D + 1], Geom<D>& G) {
    const int N = V.N;
    double x[D][D + 1];
#pragma unroll
    for (int k = 0; k < D; ++k)
#pragma unroll
        for (int a = 0; a <= D; ++a) x[k][a] = V.pos[(long long)k * N + nd[a]];
    double J[D][D];
#pragma unroll
    for (int r = 0; r < D; ++r)
#pragma unroll
        for (int c = 0; c < D; ++c) J[r][c] = x[r][c + 1] - x[r][0];
    double dj;
    if constexpr (D == 2) {
        dj = J[0][0] * J[1][1] - J[1][0] * J[0][1];
        const double inv[2][2] = {{J[1][1] / dj, -J[0][1] / dj}, {-J[1][0] / dj, J[0][0] / dj}};
#pragma unroll
        for (int k = 0; k < 2; ++k) {
            G.gN[k][0] = -inv[0][k] - inv[1][k];
            G.gN[k][1] = inv[0][k];
            G.gN[k][2] = inv[1][k];
        }
        G.vol = dj * 0.5;
    } else {
        dj = J[0][0] * J[1][1] * J[2][2] + J[0][1] * J[1][2] * J[2][0] + J[0][2] * J[1][0] * J[2][1] -
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
        G.vol = dj * 0.16666666666666666666666666666667;
    }
    const double h = sqrt((D == 2 ? 0.5 : 0.16666666666666666666666666666667) * dj / M_PI);
    double U = 0;
#pragma unroll
    for (int a = 0; a <= D; ++a) {
        double u2 = 0;
#pragma unroll
        for (int k = 0; k < D; ++k) {
            const double vi = V.q_iter[(long long)k * N + nd[a]];
            u2 += vi * vi;
        }
        U += sqrt(u2);
    }
    U /= (D + 1);
    const double dt = V.dt;
    const double visc = 4 * V.mu / (h * h * V.rho);
    G.tau = 1 / sqrt((2 / dt) * (2 / dt) + (2 * U / h) * (2 * U / h) + 9 * visc * visc);
}

// Adds the (a, b) block of Ae (row node a, column node b) to blk[s][t].
template <int D>
__device__ __forceinline__ void add_block(const PspgView& V, const Geom<D>& G, int a, int b, bool mom,
                                          double (&blk)[D + 1][D + 1]) {
    constexpr int NE = D + 1;
    const double vol = G.vol, vol_ne = vol / NE, mu_vol = V.mu * vol, tau_dt = G.tau / V.dt;
    double gdot = 0;
#pragma unroll
    for (int k = 0; k < D; ++k) gdot += G.gN[k][a] * G.gN[k][b];
    if (mom) {
        const double mab = V.rho * vol * ((a == b) ? 2.0 : 1.0) / (double)((D + 1) * (D + 2)) / V.dt;
#pragma unroll
        for (int s = 0; s < D; ++s) {
#pragma unroll
            for (int t = 0; t < D; ++t) {
                double val = mu_vol * G.gN[t][a] * G.gN[s][b];
                if (s == t) val += mab + mu_vol * gdot;
                blk[s][t] += val;
            }
            blk[s][D] += -vol_ne * G.gN[s][a];
        }
    }
#pragma unroll
    for (int t = 0; t < D; ++t) blk[D][t] += vol_ne * (tau_dt * G.gN[t][a] + G.gN[t][b]);
    blk[D][D] += G.tau / V.rho * vol * gdot;
}

__device__ __forceinline__ void decode(uint32_t w, int& e, int& a, int& b) {
    e = (int)(w >> 4);
    a = (int)((w >> 2) & 3u);
    b = (int)(w & 3u);
}

// Off-diagonal entries: lane k of node i's group owns pair (i, nbr_k), nbr_k != i.
template <int D>
__global__ __launch_bounds__(kBlock) void k_pspg_offdiag(const PspgView* __restrict__ views,
                                                         const BlockTask* __restrict__ tasks) {
    constexpr int NE = D + 1, L = kAsmLanes, NPB = kBlock / L;
    const BlockTask task = tasks[blockIdx.x];
    if (task.case_id < 0) return;
    const PspgView& V = views[task.case_id];
    const int lane = threadIdx.x % L;
    for (long long i = (long long)task.block * NPB + threadIdx.x / L; i < V.N; i += (long long)task.nblocks * NPB) {
        const int4 inf = V.node_info[i];
        const int deg = inf.z;
        const uint8_t f = (uint8_t)inf.w;
        const bool mom = (f & (NODE_BOUND | NODE_FREE)) == 0;
        const int nb = V.nbr_ptr[i];
        for (int k = lane; k < deg; k += L) {
            if (V.nbr_idx[nb + k] == (int)i) continue;
            double blk[NE][NE];
#pragma unroll
            for (int s = 0; s < NE; ++s)
#pragma unroll
                for (int t = 0; t < NE; ++t) blk[s][t] = 0.0;
            const int m = nb + k;
            for (int q = V.pei_ptr[m]; q < V.pei_ptr[m + 1]; ++q) {
                int e, a, b;
                decode(V.pei[q], e, a, b);
                int nd[NE];
#pragma unroll
                for (int c = 0; c < NE; ++c) nd[c] = V.conn[(long long)c * V.E + e];
                Geom<D> G;
                element_eval<D>(V, nd, G);
                add_block<D>(V, G, a, b, mom, blk);
            }
#pragma unroll
            for (int s = 0; s < NE; ++s) {
                if (s < D && !mom) continue;
                const long long rs = (s < D) ? (long long)s * V.mom_nnz + inf.x : (long long)D * V.mom_nnz + inf.y;
#pragma unroll
                for (int t = 0; t < NE; ++t) V.vals[rs + (long long)t * deg + k] = blk[s][t];
            }
        }
    }
}

// Diagonal block and right-hand side of node i: the node's elements (its diagonal
// pair's incidence list) are spread over the L lanes and reduced with shuffles.
template <int D>
__global__ __launch_bounds__(kBlock) void k_pspg_diag(const PspgView* __restrict__ views,
                                                      const BlockTask* __restrict__ tasks) {
    constexpr int NE = D + 1, L = kAsmLanes, NPB = kBlock / L;
    const BlockTask task = tasks[blockIdx.x];
    if (task.case_id < 0) return;
    const PspgView& V = views[task.case_id];
    const int lane = threadIdx.x % L;
    const int N = V.N;
    for (long long i = (long long)task.block * NPB + threadIdx.x / L; i < N; i += (long long)task.nblocks * NPB) {
        const int4 inf = V.node_info[i];
        const int deg = inf.z;
        const uint8_t f = (uint8_t)inf.w;
        if (f & NODE_FREE) continue;  // no elements: rows and b set by K2
        const bool mom = (f & (NODE_BOUND | NODE_FREE)) == 0;
        const int nb = V.nbr_ptr[i];
        int self = 0;  // rank of i in its own neighbourhood (sorted)
        for (int k = 0; k < deg; ++k)
            if (V.nbr_idx[nb + k] == (int)i) self = k;
        const int m = nb + self;
        double blk[NE][NE], bv[NE];
#pragma unroll
        for (int s = 0; s < NE; ++s) {
            bv[s] = 0.0;
#pragma unroll
            for (int t = 0; t < NE; ++t) blk[s][t] = 0.0;
        }
        const int q0 = V.pei_ptr[m], q1 = V.pei_ptr[m + 1];
        for (int q = q0 + lane; q < q1; q += L) {
            int e, a, b;
            decode(V.pei[q], e, a, b);
            int nd[NE];
#pragma unroll
            for (int c = 0; c < NE; ++c) nd[c] = V.conn[(long long)c * V.E + e];
            Geom<D> G;
            element_eval<D>(V, nd, G);
            add_block<D>(V, G, a, a, mom, blk);
            // be (MomContEquationPSPG.inl:53): F + Me/dt vPrev ; H + C tau/dt vPrev
            const double vol = G.vol, vol_ne = vol / NE;
            double hg = 0, cv = 0;
#pragma unroll
            for (int s = 0; s < D; ++s) {
                double mv = 0, vs = 0;
#pragma unroll
                for (int c = 0; c < NE; ++c) {
                    const double vp = V.q_prev[(long long)s * N + nd[c]];
                    mv += ((c == a) ? 2.0 : 1.0) * vp;
                    vs += vp;
                }
                bv[s] += V.rho * vol_ne * V.g[s] + V.rho * vol / (double)((D + 1) * (D + 2)) / V.dt * mv;
                hg += G.gN[s][a] * V.g[s];
                cv += G.gN[s][a] * vs;
            }
            bv[D] += G.tau * vol * hg + G.tau / V.dt * vol_ne * cv;
        }
#pragma unroll
        for (int s = 0; s < NE; ++s) {
            bv[s] = group_sum<L>(bv[s]);
#pragma unroll
            for (int t = 0; t < NE; ++t) blk[s][t] = group_sum<L>(blk[s][t]);
        }
        if (lane == 0) {
#pragma unroll
            for (int s = 0; s < NE; ++s) {
                V.b[(long long)s * N + i] = bv[s];
                if (s < D && !mom) continue;
                const long long rs = (s < D) ? (long long)s * V.mom_nnz + inf.x : (long long)D * V.mom_nnz + inf.y;
#pragma unroll
                for (int t = 0; t < NE; ++t) V.vals[rs + (long long)t * deg + self] = blk[s][t];
            }
        }
    }
}

// K2 (m_applyBCPSPG, MomContEquationPSPG.inl:190-231) plus the unit diagonals of
// MomContEquationPSPG.inl:108-128; one thread per node j.
__global__ __launch_bounds__(kBlock) void k_pspg_bc(const PspgView* __restrict__ views,
                                                    const BlockTask* __restrict__ tasks) {
    const BlockTask task = tasks[blockIdx.x];
    if (task.case_id < 0) return;
    const PspgView& V = views[task.case_id];
    const int N = V.N, D = V.dim;
    for (int j = task.block * kBlock + threadIdx.x; j < N; j += task.nblocks * kBlock) {
        const int4 inf = V.node_info[j];
        const uint8_t f = (uint8_t)inf.w;
        const bool isfree = f & NODE_FREE, bound = f & NODE_BOUND;
        const bool mom = !(bound || isfree), cont = !isfree;
        if (bound || isfree)
            for (int s = 0; s < D; ++s) V.vals[(long long)s * V.mom_nnz + inf.x] = 1.0;
        if (isfree) {
            V.vals[(long long)D * V.mom_nnz + inf.y] = 1.0;
            V.b[(long long)D * N + j] = 0.0;
            if (!bound)
                for (int s = 0; s < D; ++s) V.b[(long long)s * N + j] = V.q_prev[(long long)s * N + j] + V.dt * V.g[s];
        }
        if (bound && (f & NODE_DIRICHLET))
            for (int s = 0; s < D; ++s) V.b[(long long)s * N + j] = V.dir_vals[(long long)s * N + j];
        // column elimination into the full rows of j: neighbours in increasing node order,
        // then velocity component, as the reference's node loop
        const int deg = inf.z, nb = V.nbr_ptr[j];
        for (int k = 0; k < deg; ++k) {
            const int n = V.nbr_idx[nb + k];
            const uint8_t fn = (uint8_t)V.node_info[n].w;
            if (!((fn & NODE_BOUND) && (fn & NODE_DIRICHLET) && !(fn & NODE_FREE))) continue;
            for (int d = 0; d < D; ++d) {
                const double vbc = V.dir_vals[(long long)d * N + n];
                for (int s = 0; s <= D; ++s) {
                    if (s < D ? !mom : !cont) continue;
                    if (n == j && s == d) continue;  // the diagonal stays
                    const long long rs = (s < D) ? (long long)s * V.mom_nnz + inf.x : (long long)D * V.mom_nnz + inf.y;
                    const long long pos = rs + (long long)d * deg + k;
                    V.b[(long long)s * N + j] -= V.vals[pos] * vbc;
                    V.vals[pos] = 0.0;
                }
            }
        }
    }
}

// Block-Jacobi preconditioner: the inverse of each node's (dim+1) x (dim+1) diagonal
// block A[(i,s),(i,t)] (FP64 Gauss-Jordan with partial pivoting), stored structure of
// arrays; plus the FP32 copies of A and of the inverses for the mixed-precision solve.
__global__ __launch_bounds__(kBlock) void k_pspg_precond(const PspgView* __restrict__ views,
                                                         const BlockTask* __restrict__ tasks) {
    const BlockTask task = tasks[blockIdx.x];
    if (task.case_id < 0) return;
    const PspgView& V = views[task.case_id];
    const int N = V.N, D = V.dim, nf = D + 1;
    const long long stride = (long long)task.nblocks * kBlock;
    for (long long i = (long long)task.block * kBlock + threadIdx.x; i < N; i += stride) {
        const int4 inf = V.node_info[i];
        const uint8_t f = (uint8_t)inf.w;
        const bool mom_full = (f & (NODE_BOUND | NODE_FREE)) == 0, cont_full = (f & NODE_FREE) == 0;
        const long long cont_start = (long long)D * V.mom_nnz + inf.y;
        // rank of i in its own neighbourhood, from the diagonal of the continuity row
        const int self = cont_full ? (int)(V.diag_pos[(long long)D * N + i] - cont_start - (long long)D * inf.z) : 0;
        double m[4][8];
#pragma unroll
        for (int s = 0; s < 4; ++s)
#pragma unroll
            for (int t = 0; t < 8; ++t) m[s][t] = (t - 4 == s) ? 1.0 : 0.0;
        for (int s = 0; s < nf; ++s) {
            const bool full = (s < D) ? mom_full : cont_full;
            const long long rs = (s < D) ? (long long)s * V.mom_nnz + inf.x : cont_start;
            for (int t = 0; t < nf; ++t) {
                if (full) m[s][t] = V.vals[rs + (long long)t * inf.z + self];
                else m[s][t] = (s == t) ? V.vals[rs] : 0.0;
            }
        }
        for (int s = nf; s < 4; ++s) m[s][s] = 1.0;
        // Gauss-Jordan on [M | I]
        for (int c = 0; c < 4; ++c) {
            int piv = c;
            for (int r = c + 1; r < 4; ++r)
                if (fabs(m[r][c]) > fabs(m[piv][c])) piv = r;
            if (piv != c)
                for (int k = 0; k < 8; ++k) {
                    const double tmp = m[c][k];
                    m[c][k] = m[piv][k];
                    m[piv][k] = tmp;
                }
            const double inv = 1.0 / m[c][c];
            for (int k = 0; k < 8; ++k) m[c][k] *= inv;
            for (int r = 0; r < 4; ++r) {
                if (r == c) continue;
                const double fct = m[r][c];
                if (fct != 0.0)
                    for (int k = 0; k < 8; ++k) m[r][k] -= fct * m[c][k];
            }
        }
        for (int s = 0; s < nf; ++s)
            for (int t = 0; t < nf; ++t) {
                const double v = m[s][4 + t];
                V.binv[(long long)(s * nf + t) * N + i] = v;
                if (V.binv32) V.binv32[(long long)(s * nf + t) * N + i] = (float)v;
            }
    }
    if (V.vals32)
        for (long long k = (long long)task.block * kBlock + threadIdx.x; k < V.nnz; k += stride)
            V.vals32[k] = (float)V.vals[k];
}

// Launchers ---------------------------------------------------------------------------
void launch_pspg_assemble(int dim, const PspgView* views, const BlockTask* tasks, int ntasks, hipStream_t s) {
    if (dim == 2) {
        k_pspg_offdiag<2><<<ntasks, kBlock, 0, s>>>(views, tasks);
        k_pspg_diag<2><<<ntasks, kBlock, 0, s>>>(views, tasks);
    } else {
        k_pspg_offdiag<3><<<ntasks, kBlock, 0, s>>>(views, tasks);
        k_pspg_diag<3><<<ntasks, kBlock, 0, s>>>(views, tasks);
    }
}

void launch_pspg_bc(const PspgView* views, const BlockTask* tasks, int ntasks, hipStream_t s) {
    k_pspg_bc<<<ntasks, kBlock, 0, s>>>(views, tasks);
}

void launch_pspg_precond(const PspgView* views, const BlockTask* tasks, int ntasks, hipStream_t s) {
    k_pspg_precond<<<ntasks, kBlock, 0, s>>>(views, tasks);
}

}  // namespace pfem
