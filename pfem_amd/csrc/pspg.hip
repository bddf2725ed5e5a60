// K1 (assembly) and K2 (boundary conditions) of the PSPG momentum-continuity system,
// MomContEquationPSPG.inl:7-235, for a batch of cases.
//
// K1 is a gather assembly: every stored value is computed by the lane that owns it and
// written exactly once, with no atomics.  For node-graph entry (i, j) the elements that
// contain both nodes come from the pair-element incidence built with the pattern (sorted
// by element index, so the element contributions of an entry are summed in the same
// order as Eigen's duplicate sum of the reference's triplet list).  Each element's
// geometry (J, detJ, inverse J; Element.cpp:15-135), its P1 shape-function gradients
// (MatricesBuilder.inl:93-128) and tau (MomContEquationPSPG.inl:238-259) are evaluated
// once per assembly by an element pass into a record the gathers read; the entries of Ae
// (MomContEquationPSPG.inl:42)
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

// Element records: one pass evaluates every element once per assembly and leaves, for the
// pair and diagonal gathers, kRecW<D> doubles per element (128 B in 3D, 64 B in 2D): the
// shape-function gradients of local nodes 1..D (the columns of inv J), |e|, tau and the sums
// over the element's nodes of each previous-step velocity component.  The gathers rebuild the
// gradient of local node 0 with the same operations as element_eval, so the values equal
// those of an evaluation in place.
template <int D>
constexpr int kRecW = D == 2 ? 8 : 16;

template <int D>
__device__ __forceinline__ void load_rec(const double* __restrict__ rec, Geom<D>& G, double (&sv)[D]) {
    if constexpr (D == 2) {
        const double4 r0 = *reinterpret_cast<const double4*>(rec);
        const double4 r1 = *reinterpret_cast<const double4*>(rec + 4);
        G.gN[0][1] = r0.x;
        G.gN[1][1] = r0.y;
        G.gN[0][2] = r0.z;
        G.gN[1][2] = r0.w;
        G.vol = r1.x;
        G.tau = r1.y;
        sv[0] = r1.z;
        sv[1] = r1.w;
    } else {
        const double4 r0 = *reinterpret_cast<const double4*>(rec);
        const double4 r1 = *reinterpret_cast<const double4*>(rec + 4);
        const double4 r2 = *reinterpret_cast<const double4*>(rec + 8);
        const double2 r3 = *reinterpret_cast<const double2*>(rec + 12);
        G.gN[0][1] = r0.x;
        G.gN[1][1] = r0.y;
        G.gN[2][1] = r0.z;
        G.gN[0][2] = r0.w;
        G.gN[1][2] = r1.x;
        G.gN[2][2] = r1.y;
        G.gN[0][3] = r1.z;
        G.gN[1][3] = r1.w;
        G.gN[2][3] = r2.x;
        G.vol = r2.y;
        G.tau = r2.z;
        sv[0] = r2.w;
        sv[1] = r3.x;
        sv[2] = r3.y;
    }
#pragma unroll
    for (int k = 0; k < D; ++k) {
        double g0 = -G.gN[k][1];
#pragma unroll
        for (int a = 2; a <= D; ++a) g0 -= G.gN[k][a];
        G.gN[k][0] = g0;
    }
}

// One lane per element: the element record (see kRecW).
template <int D>
__global__ __launch_bounds__(kBlock) void k_pspg_elem(const PspgView* __restrict__ views,
                                                      const BlockTask* __restrict__ tasks) {
    constexpr int NE = D + 1;
    const BlockTask task = tasks[blockIdx.x];
    if (task.case_id < 0) return;
    const PspgView& V = views[task.case_id];
    const int N = V.N;
    const Seg sg = xcd_segment(task, V.E);
    const long long eend = (long long)sg.n0 + sg.ns;
    for (long long e = sg.n0 + (long long)sg.lb * kBlock + threadIdx.x; e < eend; e += (long long)sg.cnt * kBlock) {
        int nd[NE];
#pragma unroll
        for (int c = 0; c < NE; ++c) nd[c] = V.conn[(long long)c * V.E + e];
        Geom<D> G;
        element_eval<D>(V, nd, G);
        double sv[D];
#pragma unroll
        for (int s = 0; s < D; ++s) {
            double vs = 0;
#pragma unroll
            for (int c = 0; c < NE; ++c) vs += V.q_prev[(long long)s * N + nd[c]];
            sv[s] = vs;
        }
        double* rec = V.erec + e * kRecW<D>;
        if constexpr (D == 2) {
            *reinterpret_cast<double4*>(rec) = make_double4(G.gN[0][1], G.gN[1][1], G.gN[0][2], G.gN[1][2]);
            *reinterpret_cast<double4*>(rec + 4) = make_double4(G.vol, G.tau, sv[0], sv[1]);
        } else {
            *reinterpret_cast<double4*>(rec) = make_double4(G.gN[0][1], G.gN[1][1], G.gN[2][1], G.gN[0][2]);
            *reinterpret_cast<double4*>(rec + 4) = make_double4(G.gN[1][2], G.gN[2][2], G.gN[0][3], G.gN[1][3]);
            *reinterpret_cast<double4*>(rec + 8) = make_double4(G.gN[2][3], G.vol, G.tau, sv[0]);
            *reinterpret_cast<double4*>(rec + 12) = make_double4(sv[1], sv[2], 0.0, 0.0);
        }
    }
}

// Adds the (a, b) block of Ae (row node a, column node b) to blk[s][t].
// Column c of the gradients for a local node index c known only at run time, as a one-hot
// weighted sum (exact: one term is 1 x g, the others add zeros).  Indexing the register
// array with c, or a select chain the compiler folds back into such an index, places the
// 3D array in scratch memory, which held the assembly gathers at ~1 ms on the 262k-node dam
// break.
template <int D>
__device__ __forceinline__ void grad_col(const Geom<D>& G, int c, double (&g)[D]) {
    double w[D + 1];
#pragma unroll
    for (int a = 0; a <= D; ++a) w[a] = (a == c) ? 1.0 : 0.0;
#pragma unroll
    for (int k = 0; k < D; ++k) {
        double r = 0.0;
#pragma unroll
        for (int a = 0; a <= D; ++a) r = fma(w[a], G.gN[k][a], r);
        g[k] = r;
    }
}

template <int D>
__device__ __forceinline__ void add_block(const PspgView& V, const Geom<D>& G, int a, int b, bool mom,
                                          double (&blk)[D + 1][D + 1]) {
    constexpr int NE = D + 1;
    const double vol = G.vol, vol_ne = vol / NE, mu_vol = V.mu * vol, tau_dt = G.tau / V.dt;
    double ga[D], gb[D];
    grad_col<D>(G, a, ga);
    grad_col<D>(G, b, gb);
    double gdot = 0;
#pragma unroll
    for (int k = 0; k < D; ++k) gdot += ga[k] * gb[k];
    if (mom) {
        const double mab = V.rho * vol * ((a == b) ? 2.0 : 1.0) / (double)((D + 1) * (D + 2)) / V.dt;
#pragma unroll
        for (int s = 0; s < D; ++s) {
#pragma unroll
            for (int t = 0; t < D; ++t) {
                double val = mu_vol * ga[t] * gb[s];
                if (s == t) val += mab + mu_vol * gdot;
                blk[s][t] += val;
            }
            blk[s][D] += -vol_ne * ga[s];
        }
    }
#pragma unroll
    for (int t = 0; t < D; ++t) blk[D][t] += vol_ne * (tau_dt * ga[t] + gb[t]);
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
    const Seg sg = xcd_segment(task, V.N);
    const long long iend = (long long)sg.n0 + sg.ns;
    for (long long i = sg.n0 + (long long)sg.lb * NPB + threadIdx.x / L; i < iend; i += (long long)sg.cnt * NPB) {
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
                Geom<D> G;
                double sv[D];
                load_rec<D>(V.erec + (long long)e * kRecW<D>, G, sv);
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
    const Seg sg = xcd_segment(task, N);
    const long long iend = (long long)sg.n0 + sg.ns;
    for (long long i = sg.n0 + (long long)sg.lb * NPB + threadIdx.x / L; i < iend; i += (long long)sg.cnt * NPB) {
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
        double vown[D];  // node i's previous-step velocity (the a == i term of Me/dt vPrev)
#pragma unroll
        for (int s = 0; s < D; ++s) vown[s] = V.q_prev[(long long)s * N + i];
        const int q0 = V.pei_ptr[m], q1 = V.pei_ptr[m + 1];
        for (int q = q0 + lane; q < q1; q += L) {
            int e, a, b;
            decode(V.pei[q], e, a, b);
            Geom<D> G;
            double sv[D];
            load_rec<D>(V.erec + (long long)e * kRecW<D>, G, sv);
            add_block<D>(V, G, a, a, mom, blk);
            // be (MomContEquationPSPG.inl:53): F + Me/dt vPrev ; H + C tau/dt vPrev, with
            // sum_c M_ac v_c = |e| (v_a + sum_c v_c) / ((d+1)(d+2))
            const double vol = G.vol, vol_ne = vol / NE;
            double hg = 0, cv = 0;
            double ga[D];
            grad_col<D>(G, a, ga);
#pragma unroll
            for (int s = 0; s < D; ++s) {
                const double mv = vown[s] + sv[s];
                bv[s] += V.rho * vol_ne * V.g[s] + V.rho * vol / (double)((D + 1) * (D + 2)) / V.dt * mv;
                hg += ga[s] * V.g[s];
                cv += ga[s] * sv[s];
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
#pragma unroll
        for (int s = 0; s < 4; ++s) {
            if (s >= nf) {
                m[s][s] = 1.0;  // padding rows (2D): identity
                continue;
            }
            const bool full = (s < D) ? mom_full : cont_full;
            const long long rs = (s < D) ? (long long)s * V.mom_nnz + inf.x : cont_start;
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                if (t >= nf) continue;
                if (full) m[s][t] = V.vals[rs + (long long)t * inf.z + self];
                else m[s][t] = (s == t) ? V.vals[rs] : 0.0;
            }
        }
        gj_invert4<false>(m);
#pragma unroll
        for (int s = 0; s < 4; ++s)
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                if (s >= nf || t >= nf) continue;
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
        k_pspg_elem<2><<<ntasks, kBlock, 0, s>>>(views, tasks);
        k_pspg_offdiag<2><<<ntasks, kBlock, 0, s>>>(views, tasks);
        k_pspg_diag<2><<<ntasks, kBlock, 0, s>>>(views, tasks);
    } else {
        k_pspg_elem<3><<<ntasks, kBlock, 0, s>>>(views, tasks);
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
