// P1 element geometry shared by the WC kernels and the extractors: Jacobian determinant,
// size |e| = detJ * refSize (refSize 1/2 in 2D, 1/6 in 3D; Mesh.cpp:472-488, Element.cpp:159-162)
// and the shape-function gradients from the inverse Jacobian (Element.cpp:15-135).
#pragma once

#include "common.hpp"

namespace pfem {

template <int D>
struct WGeom {
    double gN[D][D + 1];
    double vol, detJ;
};

// Element geometry from the coordinates x[k][a] of its nodes.
template <int D>
__device__ __forceinline__ void wc_geom(const double (&x)[D][D + 1], WGeom<D>& G) {
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

template <int D>
__device__ __forceinline__ void wc_geometry(const double* __restrict__ pos, int N, const int (&nd)[D + 1],
                                            double (&x)[D][D + 1], WGeom<D>& G) {
#pragma unroll
    for (int k = 0; k < D; ++k)
#pragma unroll
        for (int a = 0; a <= D; ++a) x[k][a] = pos[(long long)k * N + nd[a]];
    wc_geom<D>(x, G);
}

}  // namespace pfem
