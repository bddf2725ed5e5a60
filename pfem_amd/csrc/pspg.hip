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
//   (Boussinesq: both scaled by 1 - alpha (T - Tr) under the integral, T linear on the element)
// Off-diagonal node pairs take one lane each; the diagonal block and the right-hand side
// of a node (every element around it) are spread over the node's lane group and reduced
// with cross-lane shuffles.  Momentum rows of boundary nodes are not assembled
// (MomContEquationPSPG.inl:68); K2 gives them, and every row of a free node, the unit
// diagonal of MomContEquationPSPG.inl:108-128.  The diagonal gather also writes the
// block-Jacobi inverse of the node's diagonal block as K2 leaves it (K2 changes a node's own
// block only through its flags), and every write of a value also writes the FP32 copies the
// mixed-precision modes keep, so no later pass re-reads the values.
//
// K2 is node-major and also atomic-free: the Dirichlet column elimination
// (MomContEquationPSPG.inl:216-228) is done row by row — the owner of row (j, s) walks its
// sorted neighbours and their velocity components, which subtracts into b[(j, s)] in the
// reference's order (Dirichlet node index, then component).
#include "common.hpp"
#include "fst.hpp"
#include "views.hpp"

namespace pfem {


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
// pair and diagonal gathers, 16 doubles per element (128 B, one cache line): the
// shape-function gradients of local nodes 1..D (the columns of inv J; the gathers rebuild
// node 0's as -(sum of the others), as element_eval does) and the element factors the
// closed forms below multiply them by, so that the gathers divide nothing:
//   vol_ne = |e|/(d+1), mu_vol = mu |e|, mab = rho |e| / ((d+1)(d+2)) / dt,
//   tvn = tau/dt |e|/(d+1), trv = tau/rho |e|, tv = tau |e|, rvn = rho |e|/(d+1);
// and, in a second array of 4 doubles per element, the sums over the element's nodes of
// each previous-step velocity component (the diagonal gather's rhs).
constexpr int kRecW = 16;
enum : int { R_VOLNE = 9, R_MUVOL, R_MAB, R_TVN, R_TRV, R_TV, R_RVN };

template <int D>
struct ERec {
    double gN[D][D + 1];
    double vol_ne, mu_vol, mab, tvn, trv, tv, rvn;
};

template <int D>
__device__ __forceinline__ void load_rec(const double* __restrict__ rec, ERec<D>& R) {
    const double4* r4 = reinterpret_cast<const double4*>(rec);
    const double4 r0 = r4[0], r2 = r4[2], r3 = r4[3];
    if constexpr (D == 2) {
        R.gN[0][1] = r0.x;
        R.gN[1][1] = r0.y;
        R.gN[0][2] = r0.z;
        R.gN[1][2] = r0.w;
    } else {
        const double4 r1 = r4[1];
        R.gN[0][1] = r0.x;
        R.gN[1][1] = r0.y;
        R.gN[2][1] = r0.z;
        R.gN[0][2] = r0.w;
        R.gN[1][2] = r1.x;
        R.gN[2][2] = r1.y;
        R.gN[0][3] = r1.z;
        R.gN[1][3] = r1.w;
        R.gN[2][3] = r2.x;
    }
    R.vol_ne = r2.y;
    R.mu_vol = r2.z;
    R.mab = r2.w;
    R.tvn = r3.x;
    R.trv = r3.y;
    R.tv = r3.z;
    R.rvn = r3.w;
#pragma unroll
    for (int k = 0; k < D; ++k) {
        double g0 = -R.gN[k][1];
#pragma unroll
        for (int a = 2; a <= D; ++a) g0 -= R.gN[k][a];
        R.gN[k][0] = g0;
    }
}

// K factor of an element (MomContEquation.inl:100-123): mu, or for a Bingham fluid
// mu + tau0 (1 - exp(-mReg g)) / g (tau0 mReg when g < 1e-15), with g = sqrt(V^T B^T ddev B V)
// = sqrt(2 sum_k L_kk^2 + sum_{k<m} (L_km + L_mk)^2) for the velocity gradient L of the
// element's node states (the current Picard iterate).
template <int D>
__device__ __forceinline__ double k_factor(const PspgView& V, const int (&nd)[D + 1], const Geom<D>& G) {
    if (V.problem != 1) return V.mu;  // 1: PFEM_PROBLEM_BINGHAM
    double Lg[D][D];
#pragma unroll
    for (int k = 0; k < D; ++k)
#pragma unroll
        for (int m = 0; m < D; ++m) {
            double t = 0;
#pragma unroll
            for (int a = 0; a <= D; ++a) t += G.gN[m][a] * V.q_iter[(long long)k * V.N + nd[a]];
            Lg[k][m] = t;
        }
    double q = 0;
#pragma unroll
    for (int k = 0; k < D; ++k) q += 2.0 * Lg[k][k] * Lg[k][k];
#pragma unroll
    for (int k = 0; k < D; ++k)
#pragma unroll
        for (int m = k + 1; m < D; ++m) {
            const double sh = Lg[k][m] + Lg[m][k];
            q += sh * sh;
        }
    const double gd = sqrt(q);
    const double mu_eq = gd < 1e-15 ? V.tau0 * V.mreg : V.tau0 * ((1.0 - exp(-V.mreg * gd)) / gd);
    return V.mu + mu_eq;
}

// One lane per element: the element record and the velocity sums.
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
        double sv[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int s = 0; s < D; ++s) {
            double vs = 0;
#pragma unroll
            for (int c = 0; c < NE; ++c) vs += V.q_prev[(long long)s * N + nd[c]];
            sv[s] = vs;
        }
        if (V.T) {  // Boussinesq: sum of the element's node temperatures
            double ts = 0;
#pragma unroll
            for (int c = 0; c < NE; ++c) ts += V.T[nd[c]];
            sv[3] = ts;
        }
        const double vol = G.vol, vol_ne = vol / NE;
        double rec[kRecW];
#pragma unroll
        for (int q = 0; q < kRecW; ++q) rec[q] = 0.0;
#pragma unroll
        for (int a = 1; a <= D; ++a)
#pragma unroll
            for (int k = 0; k < D; ++k) rec[(a - 1) * D + k] = G.gN[k][a];
        rec[R_VOLNE] = vol_ne;
        rec[R_MUVOL] = k_factor<D>(V, nd, G) * vol;
        rec[R_MAB] = V.rho * vol / (double)((D + 1) * (D + 2)) / V.dt;
        rec[R_TVN] = G.tau / V.dt * vol_ne;
        rec[R_TRV] = G.tau / V.rho * vol;
        rec[R_TV] = G.tau * vol;
        rec[R_RVN] = V.rho * vol_ne;
        double4* out = reinterpret_cast<double4*>(V.erec + e * kRecW);
#pragma unroll
        for (int q = 0; q < 4; ++q) out[q] = make_double4(rec[4 * q], rec[4 * q + 1], rec[4 * q + 2], rec[4 * q + 3]);
        reinterpret_cast<double4*>(V.esv)[e] = make_double4(sv[0], sv[1], sv[2], sv[3]);
    }
}

// Column c of the gradients for a local node index c known only at run time, as a one-hot
// weighted sum (exact: one term is 1 x g, the others add zeros).  Indexing the register
// array with c, or a select chain the compiler folds back into such an index, places the
// 3D array in scratch memory, which held the assembly gathers at ~1 ms on the 262k-node dam
// break.
template <int D>
__device__ __forceinline__ void grad_col(const ERec<D>& R, int c, double (&g)[D]) {
    double w[D + 1];
#pragma unroll
    for (int a = 0; a <= D; ++a) w[a] = (a == c) ? 1.0 : 0.0;
#pragma unroll
    for (int k = 0; k < D; ++k) {
        double r = 0.0;
#pragma unroll
        for (int a = 0; a <= D; ++a) r = fma(w[a], R.gN[k][a], r);
        g[k] = r;
    }
}

// Adds the (a, b) block of Ae (row node a, column node b) to blk[s][t].
template <int D>
__device__ __forceinline__ void add_block(const ERec<D>& R, int a, int b, bool mom, double (&blk)[D + 1][D + 1]) {
    double ga[D], gb[D];
    grad_col<D>(R, a, ga);
    grad_col<D>(R, b, gb);
    double gdot = 0;
#pragma unroll
    for (int k = 0; k < D; ++k) gdot += ga[k] * gb[k];
    if (mom) {
        const double mdiag = (a == b) ? 2.0 * R.mab : R.mab;
        const double dterm = mdiag + R.mu_vol * gdot;
#pragma unroll
        for (int s = 0; s < D; ++s) {
#pragma unroll
            for (int t = 0; t < D; ++t) {
                double val = R.mu_vol * ga[t] * gb[s];
                if (s == t) val += dterm;
                blk[s][t] += val;
            }
            blk[s][D] += -R.vol_ne * ga[s];
        }
    }
#pragma unroll
    for (int t = 0; t < D; ++t) blk[D][t] += R.tvn * ga[t] + R.vol_ne * gb[t];
    blk[D][D] += R.trv * gdot;
}

// Writes node i's block at rank k into the sliced ELL copy (EllMat layout: value (s, t) of
// lane r = i % 64 of slice c = i / 64 at ((soff[c] + k) nf^2 + s nf + t) 64 + r); rows that
// are not full hold zeros there (their unit diagonal comes from K2).  The nodes of a
// workgroup (kBlock / kAsmLanes) write the same 128-byte lines of each (k, s, t), so the L2
// merges the lanes' 8-byte stores into whole lines.
template <int NE>
__device__ __forceinline__ void ell_put(const PspgView& V, long long i, int k, const double (&blk)[NE][NE],
                                        bool mom, bool cont) {
    if (!V.vals_ell) return;
    const long long c = i / kEllSlice, r = i % kEllSlice;
    const long long base = ((long long)V.ell_soff[c] + k) * (NE * NE) * kEllSlice + r;
#pragma unroll
    for (int s = 0; s < NE; ++s) {
        const bool full = (s < NE - 1) ? mom : cont;
#pragma unroll
        for (int t = 0; t < NE; ++t) {
            const double v = full ? blk[s][t] : 0.0;
            const long long q = base + (long long)(s * NE + t) * kEllSlice;
            V.vals_ell[q] = v;
            if (V.vals32_ell) V.vals32_ell[q] = (float)v;
        }
    }
}

// Writes value q of the node-block layout and, when the precision mode keeps one, its FP32 copy.
__device__ __forceinline__ void put_val(const PspgView& V, long long q, double v) {
    V.vals[q] = v;
    if (V.vals32) V.vals32[q] = (float)v;
}

// Block-Jacobi inverse of node i (structure of arrays [(s nf + t)][N], FP64 and, when kept,
// FP32) from its final diagonal block m[0..nf)[0..nf) (m[r][4..7] = I on entry; padding rows
// of a 2D block hold the identity).
template <int NE>
__device__ __forceinline__ void put_binv(const PspgView& V, long long i, double (&m)[4][8]) {
    gj_invert4<false>(m);
#pragma unroll
    for (int s = 0; s < NE; ++s)
#pragma unroll
        for (int t = 0; t < NE; ++t) {
            const double v = m[s][4 + t];
            V.binv[(long long)(s * NE + t) * V.N + i] = v;
            if (V.binv32) V.binv32[(long long)(s * NE + t) * V.N + i] = (float)v;
        }
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
        const bool mom = (f & (NODE_BOUND | NODE_FREE)) == 0, cont = (f & NODE_FREE) == 0;
        const int nb = V.nbr_ptr[i];
        // ranks of the node's ELL slice past its degree: zeros (a free node's rank 0 too;
        // K2 sets its unit diagonal)
        const long long c = i / kEllSlice;
        const int W = V.vals_ell ? V.ell_soff[c + 1] - V.ell_soff[c] : 0;
        for (int k = deg + lane; k < W; k += L) {
            double zero[NE][NE];
#pragma unroll
            for (int s = 0; s < NE; ++s)
#pragma unroll
                for (int t = 0; t < NE; ++t) zero[s][t] = 0.0;
            ell_put<NE>(V, i, k, zero, false, false);
        }
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
                ERec<D> R;
                load_rec<D>(V.erec + (long long)e * kRecW, R);
                add_block<D>(R, a, b, mom, blk);
            }
#pragma unroll
            for (int s = 0; s < NE; ++s) {
                if (s < D && !mom) continue;
                const long long rs = (s < D) ? (long long)s * V.mom_nnz + inf.x : (long long)D * V.mom_nnz + inf.y;
#pragma unroll
                for (int t = 0; t < NE; ++t) put_val(V, rs + (long long)t * deg + k, blk[s][t]);
            }
            ell_put<NE>(V, i, k, blk, mom, cont);
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
        if (f & NODE_FREE) {  // no elements: rows and b set by K2, block-Jacobi inverse I
            if (lane == 0) {
                double m[4][8];
#pragma unroll
                for (int r = 0; r < 4; ++r)
#pragma unroll
                    for (int c = 0; c < 8; ++c) m[r][c] = (c == r || c - 4 == r) ? 1.0 : 0.0;
                put_binv<NE>(V, i, m);
            }
            continue;
        }
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
        const double town = V.T ? V.T[i] : 0.0;
        const int q0 = V.pei_ptr[m], q1 = V.pei_ptr[m + 1];
        for (int q = q0 + lane; q < q1; q += L) {
            int e, a, b;
            decode(V.pei[q], e, a, b);
            ERec<D> R;
            load_rec<D>(V.erec + (long long)e * kRecW, R);
            add_block<D>(R, a, a, mom, blk);
            // be (MomContEquationPSPG.inl:53): F + Me/dt vPrev ; H + C tau/dt vPrev, with
            // sum_c M_ac v_c = |e| (v_a + sum_c v_c) / ((d+1)(d+2))
            const double4 sv4 = reinterpret_cast<const double4*>(V.esv)[e];
            const double sv[3] = {sv4.x, sv4.y, sv4.z};
            double hg = 0, cv = 0;
            double ga[D];
            grad_col<D>(R, a, ga);
            if (V.T) {
                // Boussinesq (MomContEquation.inl:166-197): with T linear on the element,
                // int N_a rho (1 - alpha (T - Tr)) = rho |e| / (d+1) ((1 + alpha Tr) -
                // alpha (T_a + sum_c T_c) / (d+2)), and the H factor's mean is
                // 1 - alpha (sum_c T_c / (d+1) - Tr)
                const double ff = (1.0 + V.alpha * V.t_ref) - V.alpha * (town + sv4.w) / (D + 2);
                const double fh = 1.0 - V.alpha * (sv4.w / (D + 1) - V.t_ref);
#pragma unroll
                for (int s = 0; s < D; ++s) {
                    bv[s] += R.rvn * V.g[s] * ff + R.mab * (vown[s] + sv[s]);
                    hg += ga[s] * V.g[s];
                    cv += ga[s] * sv[s];
                }
                bv[D] += R.tv * hg * fh + R.tvn * cv;
            } else {
#pragma unroll
                for (int s = 0; s < D; ++s) {
                    bv[s] += R.rvn * V.g[s] + R.mab * (vown[s] + sv[s]);
                    hg += ga[s] * V.g[s];
                    cv += ga[s] * sv[s];
                }
                bv[D] += R.tv * hg + R.tvn * cv;
            }
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
                for (int t = 0; t < NE; ++t) put_val(V, rs + (long long)t * deg + self, blk[s][t]);
            }
            ell_put<NE>(V, i, self, blk, mom, true);
            // block-Jacobi inverse of the block as K2 leaves it: the velocity rows of a node
            // without momentum rows are unit rows, and a Dirichlet wall node's own velocity
            // columns are eliminated from its continuity row (k_pspg_bc)
            const bool elim = (f & NODE_BOUND) && (f & NODE_DIRICHLET);
            double m[4][8];
#pragma unroll
            for (int r = 0; r < 4; ++r)
#pragma unroll
                for (int c = 0; c < 8; ++c) {
                    double v = (c == r || c - 4 == r) ? 1.0 : 0.0;
                    if (r < NE && c < NE) {
                        if (r < D) v = mom ? blk[r][c] : (r == c ? 1.0 : 0.0);
                        else v = (elim && c < D) ? 0.0 : blk[r][c];
                    }
                    m[r][c] = v;
                }
            put_binv<NE>(V, i, m);
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
    const int N = V.N, D = V.dim, nf = D + 1;
    for (int j = task.block * kBlock + threadIdx.x; j < N; j += task.nblocks * kBlock) {
        const int4 inf = V.node_info[j];
        const uint8_t f = (uint8_t)inf.w;
        const bool isfree = f & NODE_FREE, bound = f & NODE_BOUND;
        const bool mom = !(bound || isfree), cont = !isfree;
        // surface tension first (MomContEquationPSPG.inl:155-186): the FST entries of the facets
        // whose element holds j, onto its momentum rows, before the rows are overwritten below
        if (V.fst.F > 0 && V.fst.ptr[j] < V.fst.ptr[j + 1]) {
            auto flag_of = [&](int n) { return (uint8_t)V.node_info[n].w; };
            if (D == 2) {
                double bb[2] = {V.b[j], V.b[(long long)N + j]};
                fst_add<2>(V.fst, V.conn, V.E, V.pos, N, j, flag_of, bb);
                V.b[j] = bb[0];
                V.b[(long long)N + j] = bb[1];
            } else {
                double bb[3] = {V.b[j], V.b[(long long)N + j], V.b[2LL * N + j]};
                fst_add<3>(V.fst, V.conn, V.E, V.pos, N, j, flag_of, bb);
                V.b[j] = bb[0];
                V.b[(long long)N + j] = bb[1];
                V.b[2LL * N + j] = bb[2];
            }
        }
        // the same entries in the ELL copy: value (s, t) of node j at rank k
        const long long ebase = V.vals_ell ? (long long)V.ell_soff[j / kEllSlice] * nf * nf * kEllSlice + j % kEllSlice : 0;
        auto ell_set = [&](int k, int s, int t, double v) {
            if (!V.vals_ell) return;
            const long long q = ebase + ((long long)k * nf * nf + s * nf + t) * kEllSlice;
            V.vals_ell[q] = v;
            if (V.vals32_ell) V.vals32_ell[q] = (float)v;
        };
        // rank of j among its own neighbours (a free node keeps its diagonal at rank 0)
        const int self = isfree ? 0
                                : (int)(V.diag_pos[(long long)D * N + j] - ((long long)D * V.mom_nnz + inf.y) -
                                        (long long)D * inf.z);
        if (bound || isfree)
            for (int s = 0; s < D; ++s) {
                put_val(V, (long long)s * V.mom_nnz + inf.x, 1.0);
                ell_set(self, s, s, 1.0);
            }
        if (isfree) {
            put_val(V, (long long)D * V.mom_nnz + inf.y, 1.0);
            ell_set(self, D, D, 1.0);
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
                    put_val(V, pos, 0.0);
                    ell_set(k, s, d, 0.0);
                }
            }
        }
    }
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


}  // namespace pfem
