// Surface tension on the device: the FST facet term of the reference (getP / getT / getFST,
// MatricesBuilder.inl:167-214 and :389-402; the facet normal of Facet::computeNormal,
// Facet.cpp:130-210) as a per-node gather.  Each node lists the facets whose element holds it,
// in facet order, and adds each facet's FST entry for its local index to its momentum rows, the
// order of the reference's facet loop (MomContEquationPSPG.inl:155-186, WCompNewton/
// MomEquation.inl:312-335).  A facet's FST is recomputed by each of its element's nodes (d + 1
// times): facets are few against elements, and the gather needs no facet pass or scratch.
#pragma once

#include "common.hpp"
#include "p1geom.hpp"
#include "views.hpp"

namespace pfem {

// Adds to b[0..D) the FST entries of node i (local index a in each listed facet's element).
template <int D, class FlagFn>
__device__ __forceinline__ void fst_add(const FstView& S, const int* conn, int E, const double* pos, int N, int i,
                                        FlagFn flag_of, double (&b)[D]) {
    constexpr int NE = D + 1;
    for (int q = S.ptr[i]; q < S.ptr[i + 1]; ++q) {
        const unsigned en = S.ent[q];
        const int f = (int)(en >> 2), a = (int)(en & 3u);
        int fn[D];
        int nfs = 0;
#pragma unroll
        for (int k = 0; k < D; ++k) {
            fn[k] = S.nodes[(size_t)f * D + k];
            nfs += (flag_of(fn[k]) & NODE_ONFS) ? 1 : 0;
        }
        if (S.all_nodes ? nfs < D : nfs == 0) continue;
        const int e = S.elem[f];
        int nd[NE];
#pragma unroll
        for (int c = 0; c < NE; ++c) nd[c] = conn[(size_t)c * E + e];
        int out = nd[0];
#pragma unroll
        for (int c = 0; c < NE; ++c) {
            bool onf = false;
#pragma unroll
            for (int k = 0; k < D; ++k) onf = onf || nd[c] == fn[k];
            if (!onf) out = nd[c];
        }
        auto P = [&](int node, int k) { return pos[(size_t)k * N + node]; };
        // facet measure and outward normal (Facet::computeDetJ, computeNormal)
        double detJ, n[3] = {0, 0, 0}, refs;
        if constexpr (D == 2) {
            const double J00 = (P(fn[1], 0) - P(fn[0], 0)) / 2, J10 = (P(fn[1], 1) - P(fn[0], 1)) / 2;
            detJ = sqrt(J00 * J00 + J10 * J10);
            double nx = P(fn[1], 1) - P(fn[0], 1), ny = P(fn[0], 0) - P(fn[1], 0);
            if (nx * (P(out, 0) - P(fn[0], 0)) + ny * (P(out, 1) - P(fn[0], 1)) > 0) {
                nx = -nx;
                ny = -ny;
            }
            const double nn = sqrt(nx * nx + ny * ny);
            n[0] = nx / nn;
            n[1] = ny / nn;
            refs = 2.0;
        } else {
            double u[3], v[3], w[3];
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                u[k] = P(fn[1], k) - P(fn[0], k);
                v[k] = P(fn[2], k) - P(fn[0], k);
                w[k] = P(out, k) - P(fn[0], k);
            }
            const double g0 = u[1] * v[2] - v[1] * u[2], g1 = u[2] * v[0] - v[2] * u[0], g2 = u[0] * v[1] - v[0] * u[1];
            detJ = sqrt(g0 * g0 + g1 * g1 + g2 * g2);
            double c[3] = {u[1] * v[2] - u[2] * v[1], u[2] * v[0] - u[0] * v[2], u[0] * v[1] - u[1] * v[0]};
            if (w[0] * c[0] + w[1] * c[1] + w[2] * c[2] > 0)
#pragma unroll
                for (int k = 0; k < 3; ++k) c[k] *= -1.0;
            const double nn = sqrt(c[0] * c[0] + c[1] * c[1] + c[2] * c[2]);
#pragma unroll
            for (int k = 0; k < 3; ++k) n[k] = c[k] / nn;
            refs = 0.5;
        }
        // P = I - n n and T(P) P in Voigt order (xx, yy, [zz,] xy[, xz, yz])
        double TP[6];
        if constexpr (D == 2) {
            const double p0 = 1 - n[0] * n[0], p1 = 1 - n[1] * n[1], p2 = -n[0] * n[1];
            TP[0] = (p0 * p0) * p0 + (p2 * p2) * p1 + (2 * p0 * p2) * p2;
            TP[1] = (p2 * p2) * p0 + (p1 * p1) * p1 + (2 * p2 * p1) * p2;
            TP[2] = (p0 * p2) * p0 + (p2 * p1) * p1 + (p0 * p1 + p2 * p2) * p2;
        } else {
            const double p[6] = {1 - n[0] * n[0], 1 - n[1] * n[1], 1 - n[2] * n[2],
                                 -n[0] * n[1],    -n[0] * n[2],    -n[1] * n[2]};
            double T[6][6];
            T[0][0] = p[0] * p[0]; T[0][1] = p[3] * p[3]; T[0][2] = p[4] * p[4];
            T[0][3] = 2 * p[0] * p[3]; T[0][5] = 2 * p[4] * p[3]; T[0][4] = 2 * p[0] * p[4];
            T[1][0] = T[0][1]; T[1][1] = p[1] * p[1]; T[1][2] = p[5] * p[5];
            T[1][3] = 2 * p[3] * p[1]; T[1][5] = 2 * p[5] * p[1]; T[1][4] = 2 * p[3] * p[5];
            T[2][0] = T[0][2]; T[2][1] = T[1][2]; T[2][2] = p[2] * p[2];
            T[2][3] = 2 * p[5] * p[4]; T[2][5] = 2 * p[5] * p[2]; T[2][4] = 2 * p[2] * p[4];
            T[3][0] = p[0] * p[3]; T[3][1] = p[3] * p[1]; T[3][2] = p[4] * p[5];
            T[3][3] = p[0] * p[1] + p[3] * p[3]; T[3][5] = p[4] * p[1] + p[5] * p[3];
            T[3][4] = p[4] * p[3] + p[0] * p[5];
            T[5][0] = p[3] * p[4]; T[5][1] = p[5] * p[1]; T[5][2] = p[5] * p[2];
            T[5][3] = p[4] * p[1] + p[5] * p[3]; T[5][5] = p[1] * p[2] + p[5] * p[5];
            T[5][4] = p[5] * p[4] + p[3] * p[2];
            T[4][0] = p[0] * p[4]; T[4][1] = p[3] * p[5]; T[4][2] = p[2] * p[4];
            T[4][3] = p[4] * p[3] + p[0] * p[5]; T[4][5] = p[5] * p[4] + p[3] * p[2];
            T[4][4] = p[2] * p[0] + p[4] * p[4];
#pragma unroll
            for (int r = 0; r < 6; ++r) {
                double t = 0;
#pragma unroll
                for (int c = 0; c < 6; ++c) t += T[r][c] * p[c];
                TP[r] = t;
            }
        }
        // the element's shape-function gradients; B^T (T P) for the rows (a, s)
        double x[D][D + 1];
        WGeom<D> G;
        wc_geometry<D>(pos, N, nd, x, G);
        double ga[D];
#pragma unroll
        for (int k = 0; k < D; ++k) {
            double r = 0.0;
#pragma unroll
            for (int c = 0; c < NE; ++c) r += (c == a) ? G.gN[k][c] : 0.0;
            ga[k] = r;
        }
        double bt[D];
        if constexpr (D == 2) {
            bt[0] = ga[0] * TP[0] + ga[1] * TP[2];
            bt[1] = ga[1] * TP[1] + ga[0] * TP[2];
        } else {
            bt[0] = ga[0] * TP[0] + ga[1] * TP[3] + ga[2] * TP[4];
            bt[1] = ga[1] * TP[1] + ga[0] * TP[3] + ga[2] * TP[5];
            bt[2] = ga[2] * TP[2] + ga[0] * TP[4] + ga[1] * TP[5];
        }
        // sum over the 3 facet Gauss points of the factor (gamma, or gamma + DgammaDT (T - Tr)
        // with T interpolated on the facet), then detJ times the reference facet size
        double w[3], Nq[3][D];
        if constexpr (D == 2) {
            const double xi[3] = {-sqrt(3.0 / 5.0), 0.0, sqrt(3.0 / 5.0)};
            w[0] = 5.0 / 18.0;
            w[1] = 8.0 / 18.0;
            w[2] = 5.0 / 18.0;
#pragma unroll
            for (int g = 0; g < 3; ++g) {
                Nq[g][0] = (1 - xi[g]) / 2;
                Nq[g][1] = (1 + xi[g]) / 2;
            }
        } else {
            const double gp[3][2] = {{1.0 / 6.0, 1.0 / 6.0}, {1.0 / 6.0, 2.0 / 3.0}, {2.0 / 3.0, 1.0 / 6.0}};
#pragma unroll
            for (int g = 0; g < 3; ++g) {
                Nq[g][0] = 1 - gp[g][0] - gp[g][1];
                Nq[g][1] = gp[g][0];
                Nq[g][2] = gp[g][1];
                w[g] = 1.0 / 3.0;
            }
        }
        double acc[D];
#pragma unroll
        for (int s = 0; s < D; ++s) acc[s] = 0.0;
#pragma unroll
        for (int g = 0; g < 3; ++g) {
            double fac = S.gamma;
            if (S.T) {
                double Tg = 0;
#pragma unroll
                for (int k = 0; k < D; ++k) Tg += Nq[g][k] * S.T[fn[k]];
                fac = S.gamma + S.dgdT * (Tg - S.t_ref);
            }
#pragma unroll
            for (int s = 0; s < D; ++s) acc[s] -= fac * bt[s] * w[g];
        }
        const double sc = detJ * refs;
#pragma unroll
        for (int s = 0; s < D; ++s) b[s] += acc[s] * sc;
    }
}

}  // namespace pfem
