// K4: weakly compressible explicit CDS_dpdt step for a batch of cases
// (WCompNewton/Solver.cpp:236-276, ContEquation.inl:123-166,304-413, MomEquation.inl:201-374)
// and the CFL time step (WCompNewton/Solver.cpp:192-234).
//
// Node state (v, a, p, rho) is stored in S (double or float); positions, element
// arithmetic and the lumped-mass / force accumulators are FP64.  Element passes run one
// element per lane and scatter their lumped contributions with FP64 atomics.  The P1
// element quantities are the reference's Gauss-rule integrals in closed form:
//   lump2(M(1))_i           = |e| / (d+1)
//   (M(1) P)_i              = |e| (P_i + sum P) / ((d+1)(d+2))
//   (D(K0 + K0' N.P) V)_i   = |e| (K0/(d+1) + K0' (P_i + sum P)/((d+1)(d+2))) div_e(V)
//   lump(M(N.rho))_i        = |e| (rho_i + sum rho) / ((d+1)(d+2))
//   F(N.rho)_(i,s)          = lump(M(N.rho))_i g_s
//   (K V)_(i,s)             = mu |e| (B^T ddev B V)_(i,s),  ddev the deviatoric operator
//   (D^T P)_(i,s)           = |e|/(d+1) d_sN_i sum P
#include "common.hpp"
#include "views.hpp"

namespace pfem {

template <int D>
struct WGeom {
    double gN[D][D + 1];
    double vol, detJ;
};

template <int D>
__device__ __forceinline__ void wc_geometry(const double* __restrict__ pos, int N, const int (&nd)[D + 1],
                                            double (&x)[D][D + 1], WGeom<D>& G) {
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

// v1/2 = v + dt/2 a (all nodes); x += dt v1/2 (non-fixed); clear the continuity accumulators.
template <typename S>
__global__ __launch_bounds__(kBlock) void k_wc_half(const WcView<S>* __restrict__ views,
                                                    const BlockTask* __restrict__ tasks) {
    const BlockTask task = tasks[blockIdx.x];
    if (task.case_id < 0) return;
    const WcView<S>& V = views[task.case_id];
    const int N = V.N;
    const double dt = V.dt;
    for (int i = task.block * kBlock + threadIdx.x; i < N; i += task.nblocks * kBlock) {
        const bool fixed = V.flags[i] & NODE_FIXED;
        for (int k = 0; k < V.dim; ++k) {
            const long long q = (long long)k * N + i;
            const double vh = (double)V.v[q] + 0.5 * dt * (double)V.a[q];
            V.v[q] = (S)vh;
            if (!fixed) V.pos[q] += vh * dt;
        }
        V.m_cont[i] = 0.0;
        V.f_cont[i] = 0.0;
    }
}

// Continuity (m_buildSystemdpdt, ContEquation.inl:353-413).
template <int D, typename S>
__global__ __launch_bounds__(kBlock) void k_wc_cont_elem(const WcView<S>* __restrict__ views,
                                                         const BlockTask* __restrict__ tasks) {
    constexpr int NE = D + 1;
    const BlockTask task = tasks[blockIdx.x];
    if (task.case_id < 0) return;
    const WcView<S>& V = views[task.case_id];
    const int N = V.N, E = V.E;
    const double dt = V.dt;
    const double c2 = 1.0 / (double)((D + 1) * (D + 2));
    for (long long e = (long long)task.block * kBlock + threadIdx.x; e < E; e += (long long)task.nblocks * kBlock) {
        int nd[NE];
#pragma unroll
        for (int a = 0; a < NE; ++a) nd[a] = V.conn[(long long)a * E + e];
        double x[D][NE];
        WGeom<D> G;
        wc_geometry<D>(V.pos, N, nd, x, G);
        double P[NE], psum = 0, divv = 0;
#pragma unroll
        for (int a = 0; a < NE; ++a) {
            P[a] = (double)V.p[nd[a]];
            psum += P[a];
#pragma unroll
            for (int k = 0; k < D; ++k) divv += G.gN[k][a] * (double)V.v[(long long)k * N + nd[a]];
        }
        const double vol = G.vol;
#pragma unroll
        for (int a = 0; a < NE; ++a) {
            const double ci = V.K0 / NE + V.K0p * (P[a] + psum) * c2;
            const double ml = vol / NE;
            const double mp = V.meduri ? vol * (P[a] + psum) * c2 : ml * P[a];
            atomicAdd(&V.m_cont[nd[a]], ml);
            atomicAdd(&V.f_cont[nd[a]], -dt * vol * ci * divv + mp);
        }
    }
}

// p = M^-1 F0 (free nodes: 0), rho from the Tait-Murnaghan law (ContEquation.inl:139-147,
// 319-350); clear the momentum accumulators.
template <typename S>
__global__ __launch_bounds__(kBlock) void k_wc_cont_node(const WcView<S>* __restrict__ views,
                                                         const BlockTask* __restrict__ tasks) {
    const BlockTask task = tasks[blockIdx.x];
    if (task.case_id < 0) return;
    const WcView<S>& V = views[task.case_id];
    const int N = V.N;
    for (int i = task.block * kBlock + threadIdx.x; i < N; i += task.nblocks * kBlock) {
        double p;
        if (V.flags[i] & NODE_FREE) p = 0.0;
        else p = (1.0 / V.m_cont[i]) * V.f_cont[i];
        V.p[i] = (S)p;
        V.rho[i] = (S)(pow((V.K0p / V.K0) * p + 1.0, 1.0 / V.K0p) * V.rho_star);
        V.m_mom[i] = 0.0;
        for (int k = 0; k < V.dim; ++k) V.f_mom[(long long)k * N + i] = 0.0;
    }
}

// Momentum (m_buildSystem, MomEquation.inl:229-302).
template <int D, typename S>
__global__ __launch_bounds__(kBlock) void k_wc_mom_elem(const WcView<S>* __restrict__ views,
                                                        const BlockTask* __restrict__ tasks) {
    constexpr int NE = D + 1;
    const BlockTask task = tasks[blockIdx.x];
    if (task.case_id < 0) return;
    const WcView<S>& V = views[task.case_id];
    const int N = V.N, E = V.E;
    const double c2 = 1.0 / (double)((D + 1) * (D + 2));
    for (long long e = (long long)task.block * kBlock + threadIdx.x; e < E; e += (long long)task.nblocks * kBlock) {
        int nd[NE];
#pragma unroll
        for (int a = 0; a < NE; ++a) nd[a] = V.conn[(long long)a * E + e];
        double x[D][NE];
        WGeom<D> G;
        wc_geometry<D>(V.pos, N, nd, x, G);
        double vel[D][NE], R[NE], rsum = 0, psum = 0;
#pragma unroll
        for (int a = 0; a < NE; ++a) {
#pragma unroll
            for (int k = 0; k < D; ++k) vel[k][a] = (double)V.v[(long long)k * N + nd[a]];
            R[a] = (double)V.rho[nd[a]];
            rsum += R[a];
            psum += (double)V.p[nd[a]];
        }
        // velocity gradient L[k][m] = sum_a dN_a/dx_m v_k(a)
        double Lg[D][D];
#pragma unroll
        for (int k = 0; k < D; ++k)
#pragma unroll
            for (int m = 0; m < D; ++m) {
                double s = 0;
#pragma unroll
                for (int a = 0; a < NE; ++a) s += G.gN[m][a] * vel[k][a];
                Lg[k][m] = s;
            }
        // deviatoric stress (ddev of MomEquation.inl:63-83 applied to the Voigt strain rate)
        double tr = 0;
#pragma unroll
        for (int k = 0; k < D; ++k) tr += Lg[k][k];
        double sig[D][D];
#pragma unroll
        for (int k = 0; k < D; ++k)
#pragma unroll
            for (int m = 0; m < D; ++m)
                sig[k][m] = (k == m) ? (2.0 * Lg[k][k] - (2.0 / 3.0) * tr) : (Lg[k][m] + Lg[m][k]);
        const double vol = G.vol;
        const double mvol = V.mu * vol;
#pragma unroll
        for (int a = 0; a < NE; ++a) {
            const double ml = vol * (R[a] + rsum) * c2;
            atomicAdd(&V.m_mom[nd[a]], ml);
#pragma unroll
            for (int s = 0; s < D; ++s) {
                double kv = 0;
#pragma unroll
                for (int m = 0; m < D; ++m) kv += G.gN[m][a] * sig[s][m];
                const double f = -mvol * kv + vol / NE * G.gN[s][a] * psum + ml * V.g[s];
                atomicAdd(&V.f_mom[(long long)s * N + nd[a]], f);
            }
        }
    }
}

// a = M^-1 F with the BC of MomEquation.inl:341-373; v = v1/2 + dt/2 a (:216-222).
template <typename S>
__global__ __launch_bounds__(kBlock) void k_wc_mom_node(const WcView<S>* __restrict__ views,
                                                        const BlockTask* __restrict__ tasks) {
    const BlockTask task = tasks[blockIdx.x];
    if (task.case_id < 0) return;
    const WcView<S>& V = views[task.case_id];
    const int N = V.N;
    const double dt = V.dt;
    for (int i = task.block * kBlock + threadIdx.x; i < N; i += task.nblocks * kBlock) {
        const uint8_t f = V.flags[i];
        const double invm = 1.0 / V.m_mom[i];
        for (int k = 0; k < V.dim; ++k) {
            const long long q = (long long)k * N + i;
            double acc;
            if ((f & NODE_FREE) && !(f & NODE_BOUND)) acc = V.g[k];
            else if ((f & NODE_BOUND) && (f & NODE_DIRICHLET)) acc = V.dir_vals[q];
            else acc = invm * V.f_mom[q];
            V.a[q] = (S)acc;
            V.v[q] = (S)((double)V.v[q] + 0.5 * dt * acc);
        }
    }
}

// CFL: dt^2 = min_e sc^2 he^2 / max_n(u^2, c^2, 4 alpha^2/he^2), he = 2 r_in (Element.cpp:226-294).
template <int D, typename S>
__global__ __launch_bounds__(kBlock) void k_wc_cfl(const WcView<S>* __restrict__ views,
                                                   const BlockTask* __restrict__ tasks) {
    constexpr int NE = D + 1;
    const BlockTask task = tasks[blockIdx.x];
    if (task.case_id < 0) return;
    const WcView<S>& V = views[task.case_id];
    const int N = V.N, E = V.E;
    const double sc2 = V.sc2;
    double best = 1.7976931348623157e308;
    for (long long e = (long long)task.block * kBlock + threadIdx.x; e < E; e += (long long)task.nblocks * kBlock) {
        int nd[NE];
#pragma unroll
        for (int a = 0; a < NE; ++a) nd[a] = V.conn[(long long)a * E + e];
        double x[D][NE];
        WGeom<D> G;
        wc_geometry<D>(V.pos, N, nd, x, G);
        double rin;
        if constexpr (D == 2) {
            auto dist = [&](int a, int b) {
                const double dx = x[0][a] - x[0][b], dy = x[1][a] - x[1][b];
                return sqrt(dx * dx + dy * dy);
            };
            const double s = (dist(0, 1) + dist(1, 2) + dist(0, 2)) / 2;
            rin = G.vol / s;
        } else {
            auto area2 = [&](int o, int a, int b) {
                const double ax = x[0][a] - x[0][o], ay = x[1][a] - x[1][o], az = x[2][a] - x[2][o];
                const double bx = x[0][b] - x[0][o], by = x[1][b] - x[1][o], bz = x[2][b] - x[2][o];
                const double cx = ay * bz - az * by, cy = az * bx - ax * bz, cz = ax * by - ay * bx;
                return sqrt(cx * cx + cy * cy + cz * cz);
            };
            rin = 6 * G.vol / (area2(0, 1, 2) + area2(0, 3, 2) + area2(0, 1, 3) + area2(3, 1, 2));
        }
        const double he = 2 * rin;
        double maxS = 0;
#pragma unroll
        for (int a = 0; a < NE; ++a) {
            const int i = nd[a];
            const double p = (double)V.p[i], rho = (double)V.rho[i];
            const double c2 = (V.K0 + V.K0p * p) / rho;
            double u2 = 0;
#pragma unroll
            for (int k = 0; k < D; ++k) {
                const double vk = (double)V.v[(long long)k * N + i];
                u2 += vk * vk;
            }
            const double alpha = V.mu / rho;
            maxS = fmax(fmax(u2, c2), maxS);
            maxS = fmax(maxS, 4 * alpha * alpha / (he * he));
        }
        best = fmin(best, sc2 * he * he / maxS);
    }
    best = wave_min(best);
    if ((threadIdx.x & 63) == 0) atomic_min_pos_double(V.dt2, best);
}

// Launchers ---------------------------------------------------------------------------
template <typename S>
void launch_wc_step(const WcView<S>* v, const BlockTask* node_tasks, int n_node_tasks, const BlockTask* elem2,
                    int n_elem2, const BlockTask* elem3, int n_elem3, hipStream_t s) {
    k_wc_half<S><<<n_node_tasks, kBlock, 0, s>>>(v, node_tasks);
    if (n_elem2) k_wc_cont_elem<2, S><<<n_elem2, kBlock, 0, s>>>(v, elem2);
    if (n_elem3) k_wc_cont_elem<3, S><<<n_elem3, kBlock, 0, s>>>(v, elem3);
    k_wc_cont_node<S><<<n_node_tasks, kBlock, 0, s>>>(v, node_tasks);
    if (n_elem2) k_wc_mom_elem<2, S><<<n_elem2, kBlock, 0, s>>>(v, elem2);
    if (n_elem3) k_wc_mom_elem<3, S><<<n_elem3, kBlock, 0, s>>>(v, elem3);
    k_wc_mom_node<S><<<n_node_tasks, kBlock, 0, s>>>(v, node_tasks);
}

template <typename S>
void launch_wc_cfl(const WcView<S>* v, const BlockTask* elem2, int n_elem2, const BlockTask* elem3, int n_elem3,
                   hipStream_t s) {
    if (n_elem2) k_wc_cfl<2, S><<<n_elem2, kBlock, 0, s>>>(v, elem2);
    if (n_elem3) k_wc_cfl<3, S><<<n_elem3, kBlock, 0, s>>>(v, elem3);
}

template void launch_wc_step<double>(const WcView<double>*, const BlockTask*, int, const BlockTask*, int,
                                     const BlockTask*, int, hipStream_t);
template void launch_wc_step<float>(const WcView<float>*, const BlockTask*, int, const BlockTask*, int,
                                    const BlockTask*, int, hipStream_t);
template void launch_wc_cfl<double>(const WcView<double>*, const BlockTask*, int, const BlockTask*, int, hipStream_t);
template void launch_wc_cfl<float>(const WcView<float>*, const BlockTask*, int, const BlockTask*, int, hipStream_t);

}  // namespace pfem
