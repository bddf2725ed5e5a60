// K4: weakly compressible explicit CDS_dpdt step for a batch of cases
// (WCompNewton/Solver.cpp:236-276, ContEquation.inl:123-166,304-413, MomEquation.inl:201-374)
// and the CFL time step (WCompNewton/Solver.cpp:192-234).
//
// Node state (v, a, p, rho) is stored in S (double or float); positions, element
// arithmetic and the lumped-mass / force sums are FP64.  The element passes are node
// gathers: a group of kWcLanes lanes per node walks the node's incident elements (the
// node-element incidence, sorted by element), each lane evaluating its elements' share for
// that node, and a DPP group sum adds the shares (the continuity pass first evaluates each
// element once, k_wc_cont_elem, and its gather reads one 32-byte element record per incident
// element).  The gathers read each node's coordinates,
// pressure, velocity and density from one 64-byte record (WcView::rec, written by the half
// step and the continuity node pass) instead of from d+3 separate field arrays, so an
// element visit touches d+1 cache lines rather than (d+1)(2d+2).  Each element is evaluated once per node
// (d+1 times in all) instead of scattering its d+1 shares with FP64 atomics: device-scope
// FP64 atomics on gfx950 ran at ~36 G/s, which held the 3D step at 1.6 % of the HBM
// roofline; the gather needs none, and every sum has a fixed order.  The P1 element
// quantities are the reference's Gauss-rule integrals in closed form:
//   lump2(M(1))_i           = |e| / (d+1)
//   (M(1) P)_i              = |e| (P_i + sum P) / ((d+1)(d+2))
//   (D(K0 + K0' N.P) V)_i   = |e| (K0/(d+1) + K0' (P_i + sum P)/((d+1)(d+2))) div_e(V)
//   lump(M(N.rho))_i        = |e| (rho_i + sum rho) / ((d+1)(d+2))
//   F(N.rho)_(i,s)          = lump(M(N.rho))_i g_s
//   (K V)_(i,s)             = mu |e| (B^T ddev B V)_(i,s),  ddev the deviatoric operator
//   (D^T P)_(i,s)           = |e|/(d+1) d_sN_i sum P
#include "common.hpp"
#include "fst.hpp"
#include "p1geom.hpp"
#include "views.hpp"

namespace pfem {

// Node j's record {x, y, z, p, vx, vy, vz, rho}: two 32-byte loads from one 64-byte line.
template <int D>
__device__ __forceinline__ void rec_load(const double* __restrict__ rec, int j, double (&x)[D], double& p,
                                         double (&v)[D], double& rho) {
    const double4* R = reinterpret_cast<const double4*>(rec) + 2 * (long long)j;
    const double4 r0 = R[0], r1 = R[1];
    x[0] = r0.x;
    x[1] = r0.y;
    if constexpr (D == 3) x[D - 1] = r0.z;
    p = r0.w;
    v[0] = r1.x;
    v[1] = r1.y;
    if constexpr (D == 3) v[D - 1] = r1.z;
    rho = r1.w;
}

// he = 2 r_in of an element (Element::getRin, Element.cpp:226-294): 2D |e| / semi-perimeter,
// 3D 3 |e| / total face area.
template <int D>
__device__ __forceinline__ double elem_he(const double (&x)[D][D + 1], double vol) {
    double rin;
    if constexpr (D == 2) {
        auto dist = [&](int a, int b) {
            const double dx = x[0][a] - x[0][b], dy = x[1][a] - x[1][b];
            return sqrt(dx * dx + dy * dy);
        };
        const double s = (dist(0, 1) + dist(1, 2) + dist(0, 2)) / 2;
        rin = vol / s;
    } else {
        auto area2 = [&](int o, int a, int b) {
            const double ax = x[0][a] - x[0][o], ay = x[1][a] - x[1][o], az = x[2][a] - x[2][o];
            const double bx = x[0][b] - x[0][o], by = x[1][b] - x[1][o], bz = x[2][b] - x[2][o];
            const double cx = ay * bz - az * by, cy = az * bx - ax * bz, cz = ax * by - ay * bx;
            return sqrt(cx * cx + cy * cy + cz * cz);
        };
        rin = 6 * vol / (area2(0, 1, 2) + area2(0, 3, 2) + area2(0, 1, 3) + area2(3, 1, 2));
    }
    return 2 * rin;
}

// One node's share of computeNextDT (WCompNewton/Solver.cpp:192-234) through an element of
// size he: sc^2 he^2 / max(u^2, c^2, 4 alpha^2 / he^2) with c^2 = (K0 + K0' p) / rho
// (ContEquation.inl:117-121), alpha = mu / rho (MomEquation.inl:184-198).  The reference's
// element value is the smallest of its nodes' shares (x / max(a, b) = min(x / a, x / b) in
// correctly rounded arithmetic), and a share only grows with he, so a node's smallest share
// over its elements is the one through its smallest he: the minimum over nodes of that
// equals the reference's minimum over elements, bit for bit.
__device__ __forceinline__ double cfl_share(double sc2, double he, double u2, double p, double rho, double K0,
                                            double K0p, double mu) {
    const double c2 = (K0 + K0p * p) / rho;
    const double alpha = mu / rho;
    const double maxS = fmax(fmax(u2, c2), 4 * alpha * alpha / (he * he));
    return sc2 * he * he / maxS;
}

// v1/2 = v + dt/2 a (all nodes); x += dt v1/2 (non-fixed).
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
        double xr[3] = {0.0, 0.0, 0.0}, vr[3] = {0.0, 0.0, 0.0};
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            if (k >= V.dim) continue;
            const long long q = (long long)k * N + i;
            const double vh = (double)V.v[q] + 0.5 * dt * (double)V.a[q];
            const S vs = (S)vh;
            V.v[q] = vs;
            double x = V.pos[q];
            if (!fixed) {
                x += vh * dt;
                V.pos[q] = x;
            }
            xr[k] = x;
            vr[k] = (double)vs;
        }
        double4* R = reinterpret_cast<double4*>(V.rec) + 2 * (long long)i;
        R[0] = make_double4(xr[0], xr[1], xr[2], (double)V.p[i]);
        R[1] = make_double4(vr[0], vr[1], vr[2], (double)V.rho[i]);
    }
}

// The value of a[me] for a local node index me known only at run time.
template <int NE>
__device__ __forceinline__ double at_local(const double (&a)[NE], int me) {
    double r = a[0];
#pragma unroll
    for (int k = 1; k < NE; ++k)
        if (k == me) r = a[k];
    return r;
}

// Continuity (m_buildSystemdpdt, ContEquation.inl:353-413), element part: |e|, div_e(V) and
// sum P of each element, from its nodes' records, once per element.
template <int D, typename S>
__global__ __launch_bounds__(kBlock) void k_wc_cont_elem(const WcView<S>* __restrict__ views,
                                                         const BlockTask* __restrict__ tasks) {
    constexpr int NE = D + 1;
    const BlockTask task = tasks[blockIdx.x];
    if (task.case_id < 0) return;
    const WcView<S>& V = views[task.case_id];
    const int E = V.E;
    for (long long e = (long long)task.block * kBlock + threadIdx.x; e < E; e += (long long)task.nblocks * kBlock) {
        double x[D][NE], vel[D][NE], P[NE];
#pragma unroll
        for (int a = 0; a < NE; ++a) {
            double xa[D], va[D], ra;
            rec_load<D>(V.rec, V.conn[(long long)a * E + e], xa, P[a], va, ra);
#pragma unroll
            for (int k = 0; k < D; ++k) {
                x[k][a] = xa[k];
                vel[k][a] = va[k];
            }
        }
        WGeom<D> G;
        wc_geom<D>(x, G);
        double psum = 0, divv = 0;
#pragma unroll
        for (int a = 0; a < NE; ++a) {
            psum += P[a];
#pragma unroll
            for (int k = 0; k < D; ++k) divv += G.gN[k][a] * vel[k][a];
        }
        // he for the CFL time step after this step (positions do not move again in it)
        reinterpret_cast<double4*>(V.erec)[e] = make_double4(G.vol, divv, psum, elem_he<D>(x, G.vol));
    }
}

// Continuity, node part: lumped mass lump2(M(1)) and rhs F0 = -dt D(K0 + K0' N.P) V + M(1) P
// (Meduri) or lump2(M(1)) P of node i, summed over its elements' records.
template <int D, typename S>
__global__ __launch_bounds__(kBlock) void k_wc_cont_gather(const WcView<S>* __restrict__ views,
                                                           const BlockTask* __restrict__ tasks) {
    constexpr int NE = D + 1, L = kWcLanes[D - 2], NPB = kBlock / L;
    const BlockTask task = tasks[blockIdx.x];
    if (task.case_id < 0) return;
    const WcView<S>& V = views[task.case_id];
    const int N = V.N;
    const double dt = V.dt;
    const double c2 = 1.0 / (double)((D + 1) * (D + 2));
    const int lane = threadIdx.x % L;
    const double4* __restrict__ erec = reinterpret_cast<const double4*>(V.erec);
    for (long long i = (long long)task.block * NPB + threadIdx.x / L; i < N; i += (long long)task.nblocks * NPB) {
        const int q1 = V.elem_ptr[i + 1];
        const double Pm = (double)V.p[i];
        double ml = 0.0, fc = 0.0, hmin = 1.7976931348623157e308;
        for (int q = V.elem_ptr[i] + lane; q < q1; q += L) {
            const double4 er = erec[V.elem_list[q] >> 2];
            const double vol = er.x, divv = er.y, psum = er.z;
            hmin = fmin(hmin, er.w);
            const double ci = V.K0 / NE + V.K0p * (Pm + psum) * c2;
            const double mlump = vol / NE;
            const double mp = V.meduri ? vol * (Pm + psum) * c2 : mlump * Pm;
            ml += mlump;
            fc += -dt * vol * ci * divv + mp;
        }
        ml = group_sum<L>(ml);
        fc = group_sum<L>(fc);
        hmin = group_min<L>(hmin);
        if (lane == 0) {
            V.m_cont[i] = ml;
            V.f_cont[i] = fc;
            V.hmin[i] = hmin;
        }
    }
}

// p = M^-1 F0 (free nodes: 0), rho from the Tait-Murnaghan law (ContEquation.inl:139-147,
// 319-350).
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
        const S ps = (S)p, rs = (S)(pow((V.K0p / V.K0) * p + 1.0, 1.0 / V.K0p) * V.rho_star);
        V.p[i] = ps;
        V.rho[i] = rs;
        V.rec[8 * (long long)i + 3] = (double)ps;  // the momentum gather reads the new p and rho
        V.rec[8 * (long long)i + 7] = (double)rs;
    }
}

// Momentum (m_buildSystem, MomEquation.inl:229-302): lumped mass lump(M(N.rho)) and force
// F = -K V + D^T P + F(N.rho) of node i, summed over its elements.
template <int D, typename S>
__global__ __launch_bounds__(kBlock) void k_wc_mom_gather(const WcView<S>* __restrict__ views,
                                                          const BlockTask* __restrict__ tasks) {
    constexpr int NE = D + 1, L = kWcLanes[D - 2], NPB = kBlock / L;
    const BlockTask task = tasks[blockIdx.x];
    if (task.case_id < 0) return;
    const WcView<S>& V = views[task.case_id];
    const int N = V.N;
    const double c2 = 1.0 / (double)((D + 1) * (D + 2));
    const int lane = threadIdx.x % L;
    for (long long i = (long long)task.block * NPB + threadIdx.x / L; i < N; i += (long long)task.nblocks * NPB) {
        const int q1 = V.elem_ptr[i + 1];
        double mm = 0.0, fm[D];
#pragma unroll
        for (int s = 0; s < D; ++s) fm[s] = 0.0;
        for (int q = V.elem_ptr[i] + lane; q < q1; q += L) {
            const int me = (int)(V.elem_list[q] & 3u);
            const int4 en = V.elem_nodes[q];
            int nd[NE];
            nd[0] = en.x;
            nd[1] = en.y;
            nd[2] = en.z;
            if constexpr (NE > 3) nd[NE - 1] = en.w;
            double x[D][NE], vel[D][NE], R[NE], P[NE];
#pragma unroll
            for (int a = 0; a < NE; ++a) {
                double xa[D], va[D];
                rec_load<D>(V.rec, nd[a], xa, P[a], va, R[a]);
#pragma unroll
                for (int k = 0; k < D; ++k) {
                    x[k][a] = xa[k];
                    vel[k][a] = va[k];
                }
            }
            WGeom<D> G;
            wc_geom<D>(x, G);
            double rsum = 0, psum = 0;
#pragma unroll
            for (int a = 0; a < NE; ++a) {
                rsum += R[a];
                psum += P[a];
            }
            // velocity gradient L[k][m] = sum_a dN_a/dx_m v_k(a)
            double Lg[D][D];
#pragma unroll
            for (int k = 0; k < D; ++k)
#pragma unroll
                for (int m = 0; m < D; ++m) {
                    double t = 0;
#pragma unroll
                    for (int a = 0; a < NE; ++a) t += G.gN[m][a] * vel[k][a];
                    Lg[k][m] = t;
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
            double gme[D], Rm = R[0];
#pragma unroll
            for (int m = 0; m < D; ++m) gme[m] = G.gN[m][0];
#pragma unroll
            for (int a = 1; a < NE; ++a)
                if (a == me) {
                    Rm = R[a];
#pragma unroll
                    for (int m = 0; m < D; ++m) gme[m] = G.gN[m][a];
                }
            const double mlump = vol * (Rm + rsum) * c2;
            mm += mlump;
#pragma unroll
            for (int s = 0; s < D; ++s) {
                double kv = 0;
#pragma unroll
                for (int m = 0; m < D; ++m) kv += gme[m] * sig[s][m];
                fm[s] += -mvol * kv + vol / NE * gme[s] * psum + mlump * V.g[s];
            }
        }
        mm = group_sum<L>(mm);
#pragma unroll
        for (int s = 0; s < D; ++s) fm[s] = group_sum<L>(fm[s]);
        if (lane == 0) {
            V.m_mom[i] = mm;
#pragma unroll
            for (int s = 0; s < D; ++s) V.f_mom[(long long)s * N + i] = fm[s];
        }
    }
}

// a = M^-1 F with the BC of MomEquation.inl:341-373; v = v1/2 + dt/2 a (:216-222); then
// the node's share of the next step's CFL time step from the final v, p, rho and its
// smallest incident he (free nodes belong to no element and take no part).
template <typename S>
__global__ __launch_bounds__(kBlock) void k_wc_mom_node(const WcView<S>* __restrict__ views,
                                                        const BlockTask* __restrict__ tasks) {
    const BlockTask task = tasks[blockIdx.x];
    if (task.case_id < 0) return;
    const WcView<S>& V = views[task.case_id];
    const int N = V.N;
    const double dt = V.dt;
    double best = 1.7976931348623157e308;
    for (int i = task.block * kBlock + threadIdx.x; i < N; i += task.nblocks * kBlock) {
        const uint8_t f = V.flags[i];
        const double invm = 1.0 / V.m_mom[i];
        // the force with the surface-tension entries of the facets whose element holds i
        // (WCompNewton/MomEquation.inl:312-335, added before the boundary conditions)
        double fm[3] = {0.0, 0.0, 0.0};
        for (int k = 0; k < V.dim; ++k) fm[k] = V.f_mom[(long long)k * N + i];
        if (V.fst.F > 0 && V.fst.ptr[i] < V.fst.ptr[i + 1]) {
            auto flag_of = [&](int n) { return V.flags[n]; };
            if (V.dim == 2) {
                double bb[2] = {fm[0], fm[1]};
                fst_add<2>(V.fst, V.conn, V.E, V.pos, N, i, flag_of, bb);
                fm[0] = bb[0];
                fm[1] = bb[1];
            } else {
                double bb[3] = {fm[0], fm[1], fm[2]};
                fst_add<3>(V.fst, V.conn, V.E, V.pos, N, i, flag_of, bb);
                fm[0] = bb[0];
                fm[1] = bb[1];
                fm[2] = bb[2];
            }
        }
        double u2 = 0.0;
        for (int k = 0; k < V.dim; ++k) {
            const long long q = (long long)k * N + i;
            double acc;
            if ((f & NODE_FREE) && !(f & NODE_BOUND)) acc = V.g[k];
            else if ((f & NODE_BOUND) && (f & NODE_DIRICHLET)) acc = V.dir_vals[q];
            else acc = invm * fm[k];
            V.a[q] = (S)acc;
            const S vs = (S)((double)V.v[q] + 0.5 * dt * acc);
            V.v[q] = vs;
            u2 += (double)vs * (double)vs;
        }
        if (!(f & NODE_FREE))
            best = fmin(best, cfl_share(V.sc2, V.hmin[i], u2, (double)V.p[i], (double)V.rho[i], V.K0, V.K0p, V.mu));
    }
    block_min_to(V.dt2, best);
}

// CFL on its own (no step before it): dt^2 = min_e sc^2 he^2 / max_n(u^2, c^2, 4 alpha^2/he^2),
// he = 2 r_in (Element.cpp:226-294), as the minimum of every element node's share.
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
        const double he = elem_he<D>(x, G.vol);
#pragma unroll
        for (int a = 0; a < NE; ++a) {
            const int i = nd[a];
            double u2 = 0;
#pragma unroll
            for (int k = 0; k < D; ++k) {
                const double vk = (double)V.v[(long long)k * N + i];
                u2 += vk * vk;
            }
            best = fmin(best, cfl_share(sc2, he, u2, (double)V.p[i], (double)V.rho[i], V.K0, V.K0p, V.mu));
        }
    }
    block_min_to(V.dt2, best);
}

// Launchers ---------------------------------------------------------------------------
template <typename S>
void launch_wc_step(const WcView<S>* v, const BlockTask* node_tasks, int n_node_tasks, const BlockTask* gather2,
                    int n_gather2, const BlockTask* gather3, int n_gather3, const BlockTask* elem2, int n_elem2,
                    const BlockTask* elem3, int n_elem3, hipStream_t s) {
    k_wc_half<S><<<n_node_tasks, kBlock, 0, s>>>(v, node_tasks);
    if (n_elem2) k_wc_cont_elem<2, S><<<n_elem2, kBlock, 0, s>>>(v, elem2);
    if (n_elem3) k_wc_cont_elem<3, S><<<n_elem3, kBlock, 0, s>>>(v, elem3);
    if (n_gather2) k_wc_cont_gather<2, S><<<n_gather2, kBlock, 0, s>>>(v, gather2);
    if (n_gather3) k_wc_cont_gather<3, S><<<n_gather3, kBlock, 0, s>>>(v, gather3);
    k_wc_cont_node<S><<<n_node_tasks, kBlock, 0, s>>>(v, node_tasks);
    if (n_gather2) k_wc_mom_gather<2, S><<<n_gather2, kBlock, 0, s>>>(v, gather2);
    if (n_gather3) k_wc_mom_gather<3, S><<<n_gather3, kBlock, 0, s>>>(v, gather3);
    k_wc_mom_node<S><<<n_node_tasks, kBlock, 0, s>>>(v, node_tasks);
}

template <typename S>
void launch_wc_cfl(const WcView<S>* v, const BlockTask* elem2, int n_elem2, const BlockTask* elem3, int n_elem3,
                   hipStream_t s) {
    if (n_elem2) k_wc_cfl<2, S><<<n_elem2, kBlock, 0, s>>>(v, elem2);
    if (n_elem3) k_wc_cfl<3, S><<<n_elem3, kBlock, 0, s>>>(v, elem3);
}

template void launch_wc_step<double>(const WcView<double>*, const BlockTask*, int, const BlockTask*, int,
                                     const BlockTask*, int, const BlockTask*, int, const BlockTask*, int,
                                     hipStream_t);
template void launch_wc_step<float>(const WcView<float>*, const BlockTask*, int, const BlockTask*, int,
                                    const BlockTask*, int, const BlockTask*, int, const BlockTask*, int, hipStream_t);
template void launch_wc_cfl<double>(const WcView<double>*, const BlockTask*, int, const BlockTask*, int, hipStream_t);
template void launch_wc_cfl<float>(const WcView<float>*, const BlockTask*, int, const BlockTask*, int, hipStream_t);

}  // namespace pfem
