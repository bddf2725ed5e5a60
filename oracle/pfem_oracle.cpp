// pfem_oracle.cpp — TEST INFRASTRUCTURE ONLY.
//
// CPU FP64 restatement of the reference PFEM per-time-step path (ImperatorS79/PFEM):
// P1 element matrices, PSPG triplet assembly, setFromTriplets-equivalent duplicate sum,
// Dirichlet column elimination, the Picard loop, the weakly compressible CDS_dpdt step and
// the CFL time step, and the extractors' mass and min/max.  It is the parity reference for
// the HIP library and the timed CPU baseline; it is never linked into or called by the
// product path.
//
// Every function cites the reference file:line it restates (paths relative to the
// reference repository root).  Third-party arithmetic the reference delegates to Eigen
// (SparseMatrix::setFromTriplets, SparseLU) is restated from Eigen 3.3's published
// algorithm (triplets are inserted per inner vector in input order, duplicates are summed
// in that order) or replaced by a stated stand-in (dense partial-pivot LU for tiny
// systems; BiCGSTAB + ILU(0)/Jacobi as the "Krylov stand-in for Eigen SparseLU").
//
// Pins (tests/test_oracle_*.py): closed-form P1 matrices, invariants (K*rigid = 0,
// L*1 = 0, sum gradN = 0), brute-force triplet lists on 1-2 element meshes, SciPy SuperLU
// golden solutions, lumped-mass conservation and the CFL closed form.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <numeric>
#include <stdexcept>
#include <vector>
#include <omp.h>

namespace {

using clk = std::chrono::steady_clock;
static double secs(clk::time_point a, clk::time_point b) { return std::chrono::duration<double>(b - a).count(); }

enum NodeFlag : uint8_t { BOUND = 1, FREE = 2, ONFS = 4, FIXED = 8, DIRICHLET = 16 };

// --------------------------------------------------------------------------------------
// Quadrature tables: srcs/mesh/Mesh.cpp:342-530 (getGaussPoints, getGaussWeight,
// getRefElementSize, getShapeFunctions).  nGPHD = 3 (2D) / 4 (3D):
// srcs/simulation/physics/IncompNewton/MomContEquation.inl:12-23.
// --------------------------------------------------------------------------------------
template <int D> struct Quad;
template <> struct Quad<2> {
    static constexpr int ng = 3;
    static constexpr double refSize = 0.5;  // Mesh.cpp:482
    double gp[3][2] = {{1.0 / 6.0, 1.0 / 6.0}, {1.0 / 6.0, 2.0 / 3.0}, {2.0 / 3.0, 1.0 / 6.0}};  // Mesh.cpp:384-386
    double w[3] = {1.0 / 3.0, 1.0 / 3.0, 1.0 / 3.0};  // Mesh.cpp:451
    double N[3][3];
    Quad() {
        for (int g = 0; g < 3; ++g) {  // Mesh.cpp:510-513
            N[g][0] = 1 - gp[g][0] - gp[g][1];
            N[g][1] = gp[g][0];
            N[g][2] = gp[g][1];
        }
    }
};
template <> struct Quad<3> {
    static constexpr int ng = 4;
    static constexpr double refSize = 0.16666666666666666666666666666667;  // Mesh.cpp:485
    double gp[4][3] = {{0.585410196624968, 0.138196601125011, 0.138196601125011},
                       {0.138196601125011, 0.585410196624968, 0.138196601125011},
                       {0.138196601125011, 0.138196601125011, 0.585410196624968},
                       {0.138196601125011, 0.138196601125011, 0.138196601125011}};  // Mesh.cpp:402-405
    double w[4] = {0.25, 0.25, 0.25, 0.25};  // Mesh.cpp:467
    double N[4][4];
    Quad() {
        for (int g = 0; g < 4; ++g) {  // Mesh.cpp:516-520
            N[g][0] = 1 - gp[g][0] - gp[g][1] - gp[g][2];
            N[g][1] = gp[g][0];
            N[g][2] = gp[g][1];
            N[g][3] = gp[g][2];
        }
    }
};

// --------------------------------------------------------------------------------------
// Element geometry: srcs/mesh/Element.cpp:15-135 (computeJ, computeDetJ, computeInvJ).
// --------------------------------------------------------------------------------------
struct Geo {
    double J[3][3];
    double invJ[3][3];
    double detJ;
};

template <int D>
static void computeGeo(const double* pos, int64_t N, const int* nodes, Geo& g) {
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) g.J[r][c] = g.invJ[r][c] = 0;
    for (int r = 0; r < D; ++r)
        for (int c = 0; c < D; ++c) g.J[r][c] = pos[nodes[c + 1] + r * N] - pos[nodes[0] + r * N];
    const auto& J = g.J;
    if (D == 2) {
        g.detJ = J[0][0] * J[1][1] - J[1][0] * J[0][1];
        g.invJ[0][0] = J[1][1] / g.detJ;
        g.invJ[0][1] = -J[0][1] / g.detJ;
        g.invJ[1][0] = -J[1][0] / g.detJ;
        g.invJ[1][1] = J[0][0] / g.detJ;
    } else {
        g.detJ = J[0][0] * J[1][1] * J[2][2] + J[0][1] * J[1][2] * J[2][0] + J[0][2] * J[1][0] * J[2][1] -
                 J[2][0] * J[1][1] * J[0][2] - J[2][1] * J[1][2] * J[0][0] - J[2][2] * J[1][0] * J[0][1];
        const double dj = g.detJ;
        g.invJ[0][0] = (J[1][1] * J[2][2] - J[1][2] * J[2][1]) / dj;
        g.invJ[0][1] = (J[2][1] * J[0][2] - J[2][2] * J[0][1]) / dj;
        g.invJ[0][2] = (J[0][1] * J[1][2] - J[0][2] * J[1][1]) / dj;
        g.invJ[1][0] = (J[2][0] * J[1][2] - J[1][0] * J[2][2]) / dj;
        g.invJ[1][1] = (J[0][0] * J[2][2] - J[2][0] * J[0][2]) / dj;
        g.invJ[1][2] = (J[1][0] * J[0][2] - J[0][0] * J[1][2]) / dj;
        g.invJ[2][0] = (J[1][0] * J[2][1] - J[2][0] * J[1][1]) / dj;
        g.invJ[2][1] = (J[2][0] * J[0][1] - J[0][0] * J[2][1]) / dj;
        g.invJ[2][2] = (J[0][0] * J[1][1] - J[1][0] * J[0][1]) / dj;
    }
}

// Element::getRin: srcs/mesh/Element.cpp:226-294.
template <int D>
static double getRin(const double* pos, int64_t N, const int* nodes, double detJ) {
    auto X = [&](int k, int d) { return pos[nodes[k] + d * N]; };
    if (D == 2) {
        double A = detJ * Quad<2>::refSize;
        auto dist = [&](int a, int b) {
            double dx = X(a, 0) - X(b, 0), dy = X(a, 1) - X(b, 1);
            return std::sqrt(dx * dx + dy * dy);
        };
        double a = dist(0, 1), b = dist(1, 2), c = dist(0, 2);
        double s = (a + b + c) / 2;
        return A / s;
    }
    double V = detJ * Quad<3>::refSize;
    double x0 = X(0, 0), x1 = X(1, 0), x2 = X(2, 0), x3 = X(3, 0);
    double y0 = X(0, 1), y1 = X(1, 1), y2 = X(2, 1), y3 = X(3, 1);
    double z0 = X(0, 2), z1 = X(1, 2), z2 = X(2, 2), z3 = X(3, 2);
    auto nrm = [](double a, double b, double c) { return std::sqrt(a * a + b * b + c * c); };
    double n1 = nrm((y1 - y0) * (z2 - z0) - (z1 - z0) * (y2 - y0), (z1 - z0) * (x2 - x0) - (x1 - x0) * (z2 - z0),
                    (x1 - x0) * (y2 - y0) - (y1 - y0) * (x2 - x0));
    double n2 = nrm((y3 - y0) * (z2 - z0) - (z3 - z0) * (y2 - y0), (z3 - z0) * (x2 - x0) - (x3 - x0) * (z2 - z0),
                    (x3 - x0) * (y2 - y0) - (y3 - y0) * (x2 - x0));
    double n3 = nrm((y1 - y0) * (z3 - z0) - (z1 - z0) * (y3 - y0), (z1 - z0) * (x3 - x0) - (x1 - x0) * (z3 - z0),
                    (x1 - x0) * (y3 - y0) - (y1 - y0) * (x3 - x0));
    double n4 = nrm((y1 - y3) * (z2 - z3) - (z1 - z3) * (y2 - y3), (z1 - z3) * (x2 - x3) - (x1 - x3) * (z2 - z3),
                    (x1 - x3) * (y2 - y3) - (y1 - y3) * (x2 - x3));
    return 6 * V / (n1 + n2 + n3 + n4);
}

// --------------------------------------------------------------------------------------
// MatrixBuilder: srcs/simulation/matricesBuilder/MatricesBuilder.inl.
// Matrices are dense row-major arrays; ne = D+1, nv = 3 (2D) / 6 (3D) Voigt rows.
// --------------------------------------------------------------------------------------
template <int D> struct MB {
    static constexpr int ne = D + 1;
    static constexpr int nv = (D == 2) ? 3 : 6;
    static constexpr int nd = D * ne;
    Quad<D> q;

    // getGradN: MatricesBuilder.inl:93-128 — gradN(k, 0) = -sum_r invJ(r, k); gradN(k, i+1) = invJ(i, k).
    void gradN(const Geo& g, double G[D][ne]) const {
        for (int k = 0; k < D; ++k) {
            // the reference writes -a - b (- c) left to right
            double v = -g.invJ[0][k];
            for (int r = 1; r < D; ++r) v -= g.invJ[r][k];
            G[k][0] = v;
            for (int i = 0; i < D; ++i) G[k][i + 1] = g.invJ[i][k];
        }
    }
    // getB: MatricesBuilder.inl:130-165.
    void B(const double G[D][ne], double Bm[nv][nd]) const {
        for (int r = 0; r < nv; ++r)
            for (int c = 0; c < nd; ++c) Bm[r][c] = 0;
        for (int j = 0; j < ne; ++j) {
            if (D == 2) {
                Bm[0][j] = Bm[2][j + ne] = G[0][j];
                Bm[1][j + ne] = Bm[2][j] = G[1][j];
            } else {
                Bm[0][j] = Bm[3][j + ne] = Bm[4][j + 2 * ne] = G[0][j];
                Bm[1][j + ne] = Bm[3][j] = Bm[5][j + 2 * ne] = G[1][j];
                Bm[2][j + 2 * ne] = Bm[4][j] = Bm[5][j + ne] = G[D - 1][j];
            }
        }
    }
    // getM: MatricesBuilder.inl:217-230 (factor per Gauss point).
    template <class F> void M(double detJ, F fac, double Mm[ne][ne]) const {
        for (int i = 0; i < ne; ++i)
            for (int j = 0; j < ne; ++j) Mm[i][j] = 0;
        for (int g = 0; g < Quad<D>::ng; ++g) {
            double f = fac(q.N[g]);
            for (int i = 0; i < ne; ++i)
                for (int j = 0; j < ne; ++j) Mm[i][j] += f * (q.N[g][i] * q.N[g][j]) * q.w[g];
        }
        const double s = detJ * Quad<D>::refSize;
        for (int i = 0; i < ne; ++i)
            for (int j = 0; j < ne; ++j) Mm[i][j] *= s;
    }
    // getK: MatricesBuilder.inl:277-289; K = detJ*refSize*fact * B^T ddev B, fact = sum_g mu w_g.
    void K(double detJ, double mu, const double ddev[nv][nv], const double Bm[nv][nd], double Km[nd][nd]) const {
        double fact = 0;
        for (int g = 0; g < Quad<D>::ng; ++g) fact += mu * q.w[g];
        const double s = detJ * Quad<D>::refSize * fact;
        double BtD[nd][nv];
        for (int a = 0; a < nd; ++a)
            for (int c = 0; c < nv; ++c) {
                double t = 0;
                for (int r = 0; r < nv; ++r) t += Bm[r][a] * ddev[r][c];
                BtD[a][c] = t;
            }
        for (int a = 0; a < nd; ++a)
            for (int b = 0; b < nd; ++b) {
                double t = 0;
                for (int c = 0; c < nv; ++c) t += BtD[a][c] * Bm[c][b];
                Km[a][b] = s * t;
            }
    }
    // getD: MatricesBuilder.inl:293-305; D = detJ*refSize * sumWNT * m^T B.
    template <class F> void Dm(double detJ, F fac, const double Bm[nv][nd], double Dout[ne][nd]) const {
        double sumWNT[ne] = {};
        for (int g = 0; g < Quad<D>::ng; ++g) {
            double f = fac(q.N[g]);
            for (int i = 0; i < ne; ++i) sumWNT[i] += f * q.N[g][i] * q.w[g];
        }
        const double s = detJ * Quad<D>::refSize;
        double mB[nd];
        for (int c = 0; c < nd; ++c) {
            double t = 0;
            for (int r = 0; r < D; ++r) t += Bm[r][c];  // m = (1,..,1,0,..,0)
            mB[c] = t;
        }
        for (int i = 0; i < ne; ++i)
            for (int c = 0; c < nd; ++c) Dout[i][c] = (s * sumWNT[i]) * mB[c];
    }
    // getC: MatricesBuilder.inl:325-337; C = detJ*refSize * gradN^T * sum_g(f w_g Ntilde_g).
    void C(double detJ, const double G[D][ne], double Cout[ne][nd]) const {
        double sumNW[D][nd] = {};
        for (int g = 0; g < Quad<D>::ng; ++g)
            for (int k = 0; k < D; ++k)
                for (int j = 0; j < ne; ++j) sumNW[k][j + k * ne] += 1.0 * q.N[g][j] * q.w[g];
        const double s = detJ * Quad<D>::refSize;
        for (int i = 0; i < ne; ++i)
            for (int c = 0; c < nd; ++c) {
                double t = 0;
                for (int k = 0; k < D; ++k) t += G[k][i] * sumNW[k][c];
                Cout[i][c] = s * t;
            }
    }
    // getL: MatricesBuilder.inl:309-321; L = detJ*refSize*fact * gradN^T gradN.
    void L(double detJ, double factor, const double G[D][ne], double Lout[ne][ne]) const {
        double fact = 0;
        for (int g = 0; g < Quad<D>::ng; ++g) fact += factor * q.w[g];
        const double s = detJ * Quad<D>::refSize * fact;
        for (int i = 0; i < ne; ++i)
            for (int j = 0; j < ne; ++j) {
                double t = 0;
                for (int k = 0; k < D; ++k) t += G[k][i] * G[k][j];
                Lout[i][j] = s * t;
            }
    }
    // getF: MatricesBuilder.inl:341-354; F = detJ*refSize * sum_g f_g Ntilde_g^T vec w_g.
    template <class F> void Fv(double detJ, F fac, const double* vec, double Fout[nd]) const {
        for (int c = 0; c < nd; ++c) Fout[c] = 0;
        for (int g = 0; g < Quad<D>::ng; ++g) {
            double f = fac(q.N[g]);
            for (int k = 0; k < D; ++k)
                for (int j = 0; j < ne; ++j) Fout[j + k * ne] += f * (q.N[g][j] * vec[k]) * q.w[g];
        }
        const double s = detJ * Quad<D>::refSize;
        for (int c = 0; c < nd; ++c) Fout[c] *= s;
    }
    // getH: MatricesBuilder.inl:373-386; H = detJ*refSize * sum_g f_g gradN^T vec w_g.
    template <class F> void H(double detJ, F fac, const double G[D][ne], const double* vec, double Hout[ne]) const {
        for (int i = 0; i < ne; ++i) Hout[i] = 0;
        for (int g = 0; g < Quad<D>::ng; ++g) {
            const double f = fac(q.N[g]);
            for (int i = 0; i < ne; ++i) {
                double t = 0;
                for (int k = 0; k < D; ++k) t += G[k][i] * vec[k];
                Hout[i] += f * t * q.w[g];
            }
        }
        const double s = detJ * Quad<D>::refSize;
        for (int i = 0; i < ne; ++i) Hout[i] *= s;
    }
};

// ddev and m for the incompressible solver: MomContEquation.inl:73-95.
template <int D> static void ddevIncomp(double dd[MB<D>::nv][MB<D>::nv]) {
    constexpr int nv = MB<D>::nv;
    for (int r = 0; r < nv; ++r)
        for (int c = 0; c < nv; ++c) dd[r][c] = 0;
    for (int r = 0; r < D; ++r) dd[r][r] = 2;
    for (int r = D; r < nv; ++r) dd[r][r] = 1;
}
// ddev for the weakly compressible solver (deviatoric): WCompNewton/MomEquation.inl:63-83.
template <int D> static void ddevWC(double dd[MB<D>::nv][MB<D>::nv]) {
    constexpr int nv = MB<D>::nv;
    for (int r = 0; r < nv; ++r)
        for (int c = 0; c < nv; ++c) dd[r][c] = 0;
    for (int r = 0; r < D; ++r)
        for (int c = 0; c < D; ++c) dd[r][c] = (r == c) ? 4.0 / 3 : -2.0 / 3;
    for (int r = D; r < nv; ++r) dd[r][r] = 1;
}

// --------------------------------------------------------------------------------------
// Lower-dimensional (facet) quadrature and facet geometry, used by the heat equation's facet
// terms and the surface-tension term.
// --------------------------------------------------------------------------------------
template <int D> struct QuadLD;
template <> struct QuadLD<2> {  // segments on [-1, 1], refSize 2 (Mesh.cpp:479)
    static constexpr double refSize = 2.0;
    double N[3][2], w[3] = {5.0 / 18.0, 8.0 / 18.0, 5.0 / 18.0};  // Mesh.cpp:429
    QuadLD() {
        const double xi[3] = {-std::sqrt(3.0 / 5.0), 0, std::sqrt(3.0 / 5.0)};  // Mesh.cpp:356-358
        for (int g = 0; g < 3; ++g) {
            N[g][0] = (1 - xi[g]) / 2;  // Mesh.cpp:506-507
            N[g][1] = (1 + xi[g]) / 2;
        }
    }
};
template <> struct QuadLD<3> {  // triangles, refSize 0.5 (Mesh.cpp:482)
    static constexpr double refSize = 0.5;
    double N[3][3], w[3] = {1.0 / 3.0, 1.0 / 3.0, 1.0 / 3.0};  // Mesh.cpp:451
    QuadLD() {
        const double gp[3][2] = {{1.0 / 6.0, 1.0 / 6.0}, {1.0 / 6.0, 2.0 / 3.0}, {2.0 / 3.0, 1.0 / 6.0}};
        for (int g = 0; g < 3; ++g) {
            N[g][0] = 1 - gp[g][0] - gp[g][1];
            N[g][1] = gp[g][0];
            N[g][2] = gp[g][1];
        }
    }
};

// Facet::computeJ / computeDetJ (Facet.cpp:15-77) and computeNormal (Facet.cpp:130-210)
template <int D>
static void facetGeo(const double* pos, int64_t N, const int* fn, int out, double& detJ, double nrm[3]) {
    auto P = [&](int node, int k) { return pos[node + k * N]; };
    nrm[0] = nrm[1] = nrm[2] = 0;
    if (D == 2) {
        const double J00 = (P(fn[1], 0) - P(fn[0], 0)) / 2, J10 = (P(fn[1], 1) - P(fn[0], 1)) / 2;
        detJ = std::sqrt(J00 * J00 + J10 * J10);
        double nx = P(fn[1], 1) - P(fn[0], 1), ny = P(fn[0], 0) - P(fn[1], 0);
        const double vx = P(out, 0) - P(fn[0], 0), vy = P(out, 1) - P(fn[0], 1);
        if (nx * vx + ny * vy > 0) {
            nx = -nx;
            ny = -ny;
        }
        const double nn = std::sqrt(nx * nx + ny * ny);
        nrm[0] = nx / nn;
        nrm[1] = ny / nn;
    } else {
        double J[3][2];
        for (int k = 0; k < 3; ++k) {
            J[k][0] = P(fn[1], k) - P(fn[0], k);
            J[k][1] = P(fn[2], k) - P(fn[0], k);
        }
        const double g0 = J[1][0] * J[2][1] - J[1][1] * J[2][0], g1 = J[2][0] * J[0][1] - J[2][1] * J[0][0],
                     g2 = J[0][0] * J[1][1] - J[1][0] * J[0][1];
        detJ = std::sqrt(g0 * g0 + g1 * g1 + g2 * g2);
        double a[3], b[3], n[3], v[3];
        for (int k = 0; k < 3; ++k) {
            a[k] = P(fn[1], k) - P(fn[0], k);
            b[k] = P(fn[2], k) - P(fn[0], k);
            v[k] = P(out, k) - P(fn[0], k);
        }
        n[0] = a[1] * b[2] - a[2] * b[1];
        n[1] = a[2] * b[0] - a[0] * b[2];
        n[2] = a[0] * b[1] - a[1] * b[0];
        if (v[0] * n[0] + v[1] * n[1] + v[2] * n[2] > 0)
            for (double& c : n) c *= -1.0;
        const double nn = std::sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
        for (int k = 0; k < 3; ++k) nrm[k] = n[k] / nn;
    }
}

// --------------------------------------------------------------------------------------
// PSPG assembly: srcs/simulation/physics/IncompNewton/MomContEquationPSPG.inl:7-146.
// --------------------------------------------------------------------------------------
struct Triplet {
    int64_t row = 0, col = 0;  // default Triplet() = (0, 0, 0): PSPG.inl:17
    double val = 0;
};

// Material law of the momentum equation (MomContEquation.inl:33-66, 100-123): problem 0 is a
// Newtonian fluid (K factor mu), 1 "Bingham" (K factor mu + mu_eq of the element's strain
// rate, Papanastasiou-regularised with tau0 and mReg).  Set by orc_set_material; read by the
// PSPG entry points.
// Problem 2 "Boussinesq" (MomContEquation.inl:164-197): Newtonian K, the F factor
// rho (1 - alpha (T - Tr)) and the PSPG H factor 1 - alpha (T - Tr) with T = N T_e at each
// Gauss point, T the node temperatures set by orc_set_boussinesq.
struct Material {
    int problem = 0;
    double tau0 = 0, mReg = 0;
    double alpha = 0, Tr = 0;
    const double* T = nullptr;
};
static Material g_mat;
static std::vector<double> g_T;

struct Params {
    int dim;
    int64_t N, E;
    const int* conn;
    const uint8_t* flags;  // FREE recomputed from connectivity (Node.inl: isFree = no element)
    double rho, mu, g[3], dt;
    Material mat = g_mat;
};

// K factor of an element (MomContEquation.inl:100-123): mu, plus for a Bingham fluid
// mu_eq = tau0 (1 - exp(-mReg gammaDot)) / gammaDot (tau0 mReg when gammaDot < 1e-15) with
// gammaDot = sqrt(V^T B^T ddev B V), V the element's node velocities in the current node
// states (getElementVecState, StatesFromToQ.hpp:137-170) and ddev the incompressible one.
template <int D>
static double kFactor(const Params& P, const double Bm[MB<D>::nv][D * (D + 1)], const double dd[MB<D>::nv][MB<D>::nv],
                      const int* nodes, const double* states) {
    if (P.mat.problem != 1) return P.mu;
    constexpr int ne = D + 1, nd = D * ne, nv = MB<D>::nv;
    double V[nd];
    for (int d = 0; d < D; ++d)
        for (int i = 0; i < ne; ++i) V[i + d * ne] = states[nodes[i] + d * P.N];
    double BV[nv];
    for (int r = 0; r < nv; ++r) {
        double t = 0;
        for (int c = 0; c < nd; ++c) t += Bm[r][c] * V[c];
        BV[r] = t;
    }
    double q = 0;
    for (int r = 0; r < nv; ++r) {
        double t = 0;
        for (int c = 0; c < nv; ++c) t += dd[r][c] * BV[c];
        q += BV[r] * t;
    }
    const double gammaDot = std::sqrt(q);
    double muEq = P.mat.tau0;
    if (gammaDot < 1e-15)
        muEq *= P.mat.mReg;
    else
        muEq *= (1 - std::exp(-P.mat.mReg * gammaDot)) / gammaDot;
    return P.mu + muEq;
}

// m_computeTauPSPG: PSPG.inl:238-259 (node states are the current Picard iterate).
template <int D>
static double tauPSPG(const Params& P, const Geo& geo, const int* nodes, const double* qIter) {
    const double h = std::sqrt(Quad<D>::refSize * geo.detJ / M_PI);
    double U = 0;
    for (int n = 0; n < D + 1; ++n) {
        double nodeU = 0;
        for (int d = 0; d < D; ++d) {
            double v = qIter[nodes[n] + d * P.N];
            nodeU += v * v;
        }
        U += std::sqrt(nodeU);
    }
    U /= (D + 1);
    const double dt = P.dt;
    return 1 / std::sqrt((2 / dt) * (2 / dt) + (2 * U / h) * (2 * U / h) +
                         9 * (4 * P.mu / (h * h * P.rho)) * (4 * P.mu / (h * h * P.rho)));
}

// Element Ae, be: PSPG.inl:26-53.
template <int D>
static void pspgElement(const Params& P, const MB<D>& mb, const double* pos, const int* nodes, const double* qPrev,
                        const double* qIter, double Ae[(D + 1) * (D + 1)][(D + 1) * (D + 1)],
                        double be[(D + 1) * (D + 1)]) {
    constexpr int ne = D + 1, nd = D * ne, nv = MB<D>::nv;
    Geo geo;
    computeGeo<D>(pos, P.N, nodes, geo);
    const double tau = tauPSPG<D>(P, geo, nodes, qIter);
    double G[D][ne];
    mb.gradN(geo, G);
    double Bm[nv][nd];
    mb.B(G, Bm);
    double Ms[ne][ne];
    mb.M(geo.detJ, [&](const double*) { return P.rho; }, Ms);
    for (int i = 0; i < ne; ++i)
        for (int j = 0; j < ne; ++j) Ms[i][j] = (1 / P.dt) * Ms[i][j];
    double dd[nv][nv];
    ddevIncomp<D>(dd);
    double Ke[nd][nd];
    mb.K(geo.detJ, kFactor<D>(P, Bm, dd, nodes, qIter), dd, Bm, Ke);
    double De[ne][nd];
    mb.Dm(geo.detJ, [](const double*) { return 1.0; }, Bm, De);
    double Ce[ne][nd];
    mb.C(geo.detJ, G, Ce);
    for (int i = 0; i < ne; ++i)
        for (int c = 0; c < nd; ++c) Ce[i][c] = (tau / P.dt) * Ce[i][c];
    double Le[ne][ne];
    mb.L(geo.detJ, 1 / P.rho, G, Le);
    for (int i = 0; i < ne; ++i)
        for (int j = 0; j < ne; ++j) Le[i][j] = tau * Le[i][j];
    // T at a Gauss point (Boussinesq): N T_e
    auto Tg = [&](const double* Ng) {
        double t = 0;
        for (int a = 0; a < ne; ++a) t += Ng[a] * P.mat.T[nodes[a]];
        return t;
    };
    const bool bq = P.mat.problem == 2;
    double Fe[nd];
    mb.Fv(geo.detJ, [&](const double* Ng) { return bq ? P.rho * (1 - P.mat.alpha * (Tg(Ng) - P.mat.Tr)) : P.rho; },
          P.g, Fe);
    double He[ne];
    mb.H(geo.detJ, [&](const double* Ng) { return bq ? 1 - P.mat.alpha * (Tg(Ng) - P.mat.Tr) : 1.0; }, G, P.g, He);
    for (int i = 0; i < ne; ++i) He[i] = tau * He[i];

    constexpr int na = (D + 1) * ne;
    for (int r = 0; r < na; ++r)
        for (int c = 0; c < na; ++c) Ae[r][c] = 0;
    // Ae << Me_dt + Ke, -De^T, Ce_dt + De, Le   (PSPG.inl:42)
    for (int a = 0; a < nd; ++a)
        for (int b = 0; b < nd; ++b) {
            double me = ((a / ne) == (b / ne)) ? Ms[a % ne][b % ne] : 0.0;
            Ae[a][b] = me + Ke[a][b];
        }
    for (int a = 0; a < nd; ++a)
        for (int j = 0; j < ne; ++j) Ae[a][nd + j] = -De[j][a];
    for (int i = 0; i < ne; ++i)
        for (int b = 0; b < nd; ++b) Ae[nd + i][b] = Ce[i][b] + De[i][b];
    for (int i = 0; i < ne; ++i)
        for (int j = 0; j < ne; ++j) Ae[nd + i][nd + j] = Le[i][j];
    // vPrev = getElementVecState(qPrev): StatesFromToQ.hpp:137-170
    double vPrev[nd];
    for (int d = 0; d < D; ++d)
        for (int i = 0; i < ne; ++i) vPrev[i + d * ne] = qPrev[nodes[i] + d * P.N];
    // be << Fe + Me_dt*vPrev, He + Ce_dt*vPrev   (PSPG.inl:53)
    for (int a = 0; a < nd; ++a) {
        double t = 0;
        for (int b = 0; b < nd; ++b) {
            double me = ((a / ne) == (b / ne)) ? Ms[a % ne][b % ne] : 0.0;
            t += me * vPrev[b];
        }
        be[a] = Fe[a] + t;
    }
    for (int i = 0; i < ne; ++i) {
        double t = 0;
        for (int b = 0; b < nd; ++b) t += Ce[i][b] * vPrev[b];
        be[nd + i] = He[i] + t;
    }
}

struct CsrSys {
    int64_t n = 0;
    std::vector<int> rowPtr, colIdx;
    std::vector<double> vals, b;
    double tTriplets = 0, tPush = 0, tAssemble = 0, tVector = 0, tBC = 0;
};

static std::vector<uint8_t> derivedFlags(const Params& P) {
    std::vector<uint8_t> f(P.flags, P.flags + P.N);
    std::vector<char> used(P.N, 0);
    for (int64_t k = 0; k < P.E * (P.dim + 1); ++k) used[P.conn[k]] = 1;
    for (int64_t n = 0; n < P.N; ++n) f[n] = (f[n] & ~FREE) | (used[n] ? 0 : FREE);
    return f;
}

// Surface tension (MatricesBuilder.inl:167-214 getP / getT, :389-402 getFST): the boundary facets
// of the mesh (nodes in Facet::m_nodesIndexes order, the element each belongs to), gamma and
// DgammaDT, set by orc_set_facets.  Facet::getOutNode is the element's node that is not on the
// facet.  A problem other than Boussinesq reads gamma alone (its DgammaDT is not read: 0).
struct FstSet {
    int64_t F = 0;
    std::vector<int> nodes, elem;
    double gamma = 0, dgdT = 0;
};
static FstSet g_fst;

static bool fstActive() { return !(g_fst.gamma < 1e-15 && g_fst.dgdT < 1e-15); }

// getFST of facet f: -sum_g fac_g B^T T(P) P w_g, times detJ refSize(dim - 1); fac_g = gamma, or
// gamma + DgammaDT (T_g - Tr) with T_g = N_LD T_facet for a Boussinesq material
template <int D>
static void facetFST(const double* pos, int64_t N, const int* conn, int64_t f, const Material& mat,
                     double FST[D * (D + 1)]) {
    constexpr int ne = D + 1, nd = D * ne, nv = MB<D>::nv;
    const int* fn = g_fst.nodes.data() + f * D;
    const int* en = conn + (int64_t)g_fst.elem[f] * ne;
    int out = en[0];
    for (int a = 0; a < ne; ++a) {
        bool onf = false;
        for (int i = 0; i < D; ++i) onf = onf || en[a] == fn[i];
        if (!onf) out = en[a];
    }
    double detJ, n[3];
    facetGeo<D>(pos, N, fn, out, detJ, n);
    double Pv[nv], Tm[nv][nv];
    if (D == 2) {
        Pv[0] = 1 - n[0] * n[0];
        Pv[1] = 1 - n[1] * n[1];
        Pv[2] = -n[0] * n[1];
        Tm[0][0] = Pv[0] * Pv[0];
        Tm[0][1] = Pv[2] * Pv[2];
        Tm[0][2] = 2 * Pv[0] * Pv[2];
        Tm[1][0] = Tm[0][1];
        Tm[1][1] = Pv[1] * Pv[1];
        Tm[1][2] = 2 * Pv[2] * Pv[1];
        Tm[2][0] = Pv[0] * Pv[2];
        Tm[2][1] = Pv[2] * Pv[1];
        Tm[2][2] = Pv[0] * Pv[1] + Pv[2] * Pv[2];
    } else {
        // Voigt order xx, yy, zz, xy, xz, yz
        Pv[0] = 1 - n[0] * n[0];
        Pv[1] = 1 - n[1] * n[1];
        Pv[2] = 1 - n[2] * n[2];
        Pv[3] = -n[0] * n[1];
        Pv[4] = -n[0] * n[2];
        Pv[5] = -n[1] * n[2];
        const double* P = Pv;
        double (*T)[nv] = Tm;
        T[0][0] = P[0] * P[0]; T[0][1] = P[3] * P[3]; T[0][2] = P[4] * P[4];
        T[0][3] = 2 * P[0] * P[3]; T[0][5] = 2 * P[4] * P[3]; T[0][4] = 2 * P[0] * P[4];
        T[1][0] = T[0][1]; T[1][1] = P[1] * P[1]; T[1][2] = P[5] * P[5];
        T[1][3] = 2 * P[3] * P[1]; T[1][5] = 2 * P[5] * P[1]; T[1][4] = 2 * P[3] * P[5];
        T[2][0] = T[0][2]; T[2][1] = T[1][2]; T[2][2] = P[2] * P[2];
        T[2][3] = 2 * P[5] * P[4]; T[2][5] = 2 * P[5] * P[2]; T[2][4] = 2 * P[2] * P[4];
        T[3][0] = P[0] * P[3]; T[3][1] = P[3] * P[1]; T[3][2] = P[4] * P[5];
        T[3][3] = P[0] * P[1] + P[3] * P[3]; T[3][5] = P[4] * P[1] + P[5] * P[3]; T[3][4] = P[4] * P[3] + P[0] * P[5];
        T[5][0] = P[3] * P[4]; T[5][1] = P[5] * P[1]; T[5][2] = P[5] * P[2];
        T[5][3] = P[4] * P[1] + P[5] * P[3]; T[5][5] = P[1] * P[2] + P[5] * P[5]; T[5][4] = P[5] * P[4] + P[3] * P[2];
        T[4][0] = P[0] * P[4]; T[4][1] = P[3] * P[5]; T[4][2] = P[2] * P[4];
        T[4][3] = P[4] * P[3] + P[0] * P[5]; T[4][5] = P[5] * P[4] + P[3] * P[2]; T[4][4] = P[2] * P[0] + P[4] * P[4];
    }
    double TP[nv];
    for (int r = 0; r < nv; ++r) {
        double t = 0;
        for (int c = 0; c < nv; ++c) t += Tm[r][c] * Pv[c];
        TP[r] = t;
    }
    MB<D> mb;
    Geo geo;
    computeGeo<D>(pos, N, en, geo);
    double G[D][ne], Bm[nv][nd];
    mb.gradN(geo, G);
    mb.B(G, Bm);
    double BtTP[nd];
    for (int c = 0; c < nd; ++c) {
        double t = 0;
        for (int r = 0; r < nv; ++r) t += Bm[r][c] * TP[r];
        BtTP[c] = t;
    }
    QuadLD<D> q;
    for (int c = 0; c < nd; ++c) FST[c] = 0;
    for (int g = 0; g < 3; ++g) {
        double fac = g_fst.gamma;
        if (mat.problem == 2) {
            double Tg = 0;
            for (int i = 0; i < D; ++i) Tg += q.N[g][i] * mat.T[fn[i]];
            fac = g_fst.gamma + g_fst.dgdT * (Tg - mat.Tr);
        }
        for (int c = 0; c < nd; ++c) FST[c] -= fac * BtTP[c] * q.w[g];
    }
    for (int c = 0; c < nd; ++c) FST[c] *= detJ * q.refSize;
}

// The facet loop of m_applyBCPSPG (MomContEquationPSPG.inl:155-186): every facet with a node on
// the free surface (allNodes: every node, the WC test of MomEquation.inl:320-322) adds its FST to
// the momentum rows of its element's nodes, in facet order.
template <int D>
static void addFST(const double* pos, int64_t N, const int* conn, const std::vector<uint8_t>& fl, const Material& mat,
                   bool allNodes, double* b) {
    if (!fstActive()) return;
    constexpr int ne = D + 1;
    for (int64_t f = 0; f < g_fst.F; ++f) {
        const int* fn = g_fst.nodes.data() + f * D;
        int nfs = 0;
        for (int i = 0; i < D; ++i) nfs += (fl[fn[i]] & ONFS) ? 1 : 0;
        if (allNodes ? nfs < D : nfs == 0) continue;
        double FST[D * ne];
        facetFST<D>(pos, N, conn, f, mat, FST);
        const int* en = conn + (int64_t)g_fst.elem[f] * ne;
        for (int a = 0; a < ne; ++a)
            for (int d = 0; d < D; ++d) b[en[a] + d * N] += FST[a + d * ne];
    }
}

// m_buildAbPSPG: PSPG.inl:7-146, including the default-triplet padding (PSPG.inl:17), the
// serial unit-diagonal push-back (PSPG.inl:108-128) and Eigen 3.3's setFromTriplets
// (duplicates summed per inner vector in triplet order), emitted here as CSR.
template <int D>
static void buildAbPSPG(const Params& P, const std::vector<uint8_t>& fl, const double* pos, const double* qPrev,
                        const double* qIter, CsrSys& S) {
    constexpr int ne = D + 1;
    constexpr int na = (D + 1) * ne;
    const int64_t N = P.N, E = P.E;
    const int64_t tripletPerElm = (int64_t)na * na;
    const int64_t n = (D + 1) * N;
    MB<D> mb;
    auto t0 = clk::now();
    std::vector<Triplet> indexA(tripletPerElm * E);
    std::vector<std::pair<int64_t, double>> indexb(na * E);
#pragma omp parallel for schedule(static)
    for (int64_t elm = 0; elm < E; ++elm) {
        const int* nodes = P.conn + elm * ne;
        double Ae[na][na], be[na];
        pspgElement<D>(P, mb, pos, nodes, qPrev, qIter, Ae, be);
        int64_t countA = 0, countb = 0;
        for (int i = 0; i < ne; ++i) {
            const uint8_t fi = fl[nodes[i]];
            const bool boundOrFree = (fi & (BOUND | FREE)) != 0;
            const bool isFree = (fi & FREE) != 0;
            for (int j = 0; j < ne; ++j) {
                for (int d1 = 0; d1 < D; ++d1)
                    for (int d2 = 0; d2 <= D; ++d2) {
                        if (!boundOrFree)
                            indexA[tripletPerElm * elm + countA] = Triplet{nodes[i] + d1 * N, nodes[j] + d2 * N,
                                                                           Ae[i + d1 * ne][j + d2 * ne]};
                        countA++;
                    }
                for (int d2 = 0; d2 <= D; ++d2) {
                    if (!isFree)
                        indexA[tripletPerElm * elm + countA] =
                            Triplet{nodes[i] + D * N, nodes[j] + d2 * N, Ae[i + D * ne][j + d2 * ne]};
                    countA++;
                }
            }
            for (int d = 0; d <= D; ++d) {
                indexb[na * elm + countb] = {nodes[i] + d * N, be[i + d * ne]};
                countb++;
            }
        }
    }
    auto t1 = clk::now();
    for (int64_t nn = 0; nn < N; ++nn) {  // PSPG.inl:108-128
        if (fl[nn] & FREE) indexA.push_back(Triplet{nn + D * N, nn + D * N, 1});
        if (fl[nn] & (BOUND | FREE))
            for (int d = 0; d < D; ++d) indexA.push_back(Triplet{nn + d * N, nn + d * N, 1});
    }
    auto t2 = clk::now();
    // setFromTriplets (Eigen 3.3 SparseMatrix::setFromTriplets, PSPG.inl:136): count per
    // outer vector, insert in input order, collapse duplicates by summation in insertion
    // order, then sort.  Rows are the outer vectors here (CSR); for each (row, col) the sum
    // order is the triplet order, as in Eigen's column-major build (whose inner vectors are
    // also filled in triplet order before the transposed copy).
    std::vector<int64_t> cnt(n + 1, 0);
    for (const auto& t : indexA) cnt[t.row + 1]++;
    for (int64_t r = 0; r < n; ++r) cnt[r + 1] += cnt[r];
    std::vector<int64_t> cur(cnt.begin(), cnt.end() - 1);
    std::vector<int64_t> ecol(indexA.size());
    std::vector<double> eval(indexA.size());
    for (const auto& t : indexA) {
        int64_t k = cur[t.row]++;
        ecol[k] = t.col;
        eval[k] = t.val;
    }
    std::vector<int64_t>().swap(cur);
    S.n = n;
    S.rowPtr.assign(n + 1, 0);
    std::vector<int64_t> rowLen(n, 0);
    // collapse duplicates per row (collapseDuplicates with scalar_sum_op)
    std::vector<std::vector<std::pair<int64_t, double>>> rows(n);
#pragma omp parallel
    {
        std::vector<int64_t> seen;
#pragma omp for schedule(dynamic, 1024)
        for (int64_t r = 0; r < n; ++r) {
            auto& out = rows[r];
            out.clear();
            for (int64_t k = cnt[r]; k < cnt[r + 1]; ++k) {
                int64_t c = ecol[k];
                bool found = false;
                for (auto& pr : out)
                    if (pr.first == c) {
                        pr.second += eval[k];
                        found = true;
                        break;
                    }
                if (!found) out.push_back({c, eval[k]});
            }
            std::sort(out.begin(), out.end(),
                      [](const std::pair<int64_t, double>& a, const std::pair<int64_t, double>& b) {
                          return a.first < b.first;
                      });
        }
    }
    for (int64_t r = 0; r < n; ++r) S.rowPtr[r + 1] = S.rowPtr[r] + (int)rows[r].size();
    S.colIdx.resize(S.rowPtr[n]);
    S.vals.resize(S.rowPtr[n]);
#pragma omp parallel for schedule(static)
    for (int64_t r = 0; r < n; ++r) {
        int64_t k = S.rowPtr[r];
        for (const auto& pr : rows[r]) {
            S.colIdx[k] = (int)pr.first;
            S.vals[k] = pr.second;
            ++k;
        }
    }
    auto t3 = clk::now();
    // b scatter in triplet order: PSPG.inl:139-145
    S.b.assign(n, 0.0);
    for (const auto& db : indexb) S.b[db.first] += db.second;
    auto t4 = clk::now();
    S.tTriplets += secs(t0, t1);
    S.tPush += secs(t1, t2);
    S.tAssemble += secs(t2, t3);
    S.tVector += secs(t3, t4);
}

// m_applyBCPSPG: PSPG.inl:149-235.  The FST loop is skipped because gamma = 0 (checked
// by the caller); Dirichlet values are the host-evaluated "<Tag>V"(pos, t+dt) results.
template <int D>
static void applyBCPSPG(const Params& P, const std::vector<uint8_t>& fl, const double* pos, const double* qPrev,
                        const double* dirVals, CsrSys& S) {
    auto t0 = clk::now();
    const int64_t N = P.N;
    addFST<D>(pos, N, P.conn, fl, P.mat, false, S.b.data());
    // column index (CSC view) of the CSR: for each column the (row, k) entries in row order
    std::vector<int64_t> ccnt(S.n + 1, 0);
    for (size_t k = 0; k < S.colIdx.size(); ++k) ccnt[S.colIdx[k] + 1]++;
    for (int64_t c = 0; c < S.n; ++c) ccnt[c + 1] += ccnt[c];
    std::vector<int64_t> cpos(ccnt.begin(), ccnt.end() - 1);
    std::vector<int64_t> cRow(S.colIdx.size()), cK(S.colIdx.size());
    for (int64_t r = 0; r < S.n; ++r)
        for (int64_t k = S.rowPtr[r]; k < S.rowPtr[r + 1]; ++k) {
            int64_t c = S.colIdx[k];
            cRow[cpos[c]] = r;
            cK[cpos[c]] = k;
            cpos[c]++;
        }
    for (int64_t nn = 0; nn < N; ++nn) {
        const uint8_t f = fl[nn];
        if (f & FREE) {
            S.b[nn + D * N] = 0;
            if (!(f & BOUND))
                for (int d = 0; d < D; ++d) S.b[nn + d * N] = qPrev[nn + d * N] + P.dt * P.g[d];
        }
        if ((f & BOUND) && (f & DIRICHLET)) {
            for (int d = 0; d < D; ++d) {
                const double res = dirVals[nn + d * N];
                const int64_t col = nn + d * N;
                S.b[col] = res;
                for (int64_t k = ccnt[col]; k < ccnt[col + 1]; ++k) {
                    const int64_t row = cRow[k];
                    if (row == col) continue;
                    const double value = S.vals[cK[k]];
                    S.b[row] -= value * res;
                    S.vals[cK[k]] = 0;
                }
            }
        }
    }
    S.tBC += secs(t0, clk::now());
}

// --------------------------------------------------------------------------------------
// Sparse linear algebra (stand-ins for Eigen SparseLU, MomContEquation.hpp:6-12).
// --------------------------------------------------------------------------------------
static void spmv(int64_t n, const int* rp, const int* ci, const double* v, const double* x, double* y) {
#pragma omp parallel for schedule(static)
    for (int64_t r = 0; r < n; ++r) {
        double s = 0;
        for (int64_t k = rp[r]; k < rp[r + 1]; ++k) s += v[k] * x[ci[k]];
        y[r] = s;
    }
}
static double dot(int64_t n, const double* a, const double* b) {
    double s = 0;
#pragma omp parallel for reduction(+ : s) schedule(static)
    for (int64_t i = 0; i < n; ++i) s += a[i] * b[i];
    return s;
}

struct Precond {
    int kind = 0;  // 0 Jacobi, 1 ILU(0)
    int64_t n = 0;
    const int* rp = nullptr;
    const int* ci = nullptr;
    std::vector<double> dinv, lu;
    std::vector<int64_t> diag;
    void setup(int k, int64_t nn, const int* rp_, const int* ci_, const double* v) {
        kind = k;
        n = nn;
        rp = rp_;
        ci = ci_;
        diag.assign(n, -1);
        for (int64_t r = 0; r < n; ++r)
            for (int64_t q = rp[r]; q < rp[r + 1]; ++q)
                if (ci[q] == r) diag[r] = q;
        if (kind == 0) {
            dinv.resize(n);
            for (int64_t r = 0; r < n; ++r) dinv[r] = 1.0 / v[diag[r]];
            return;
        }
        // ILU(0) (IKJ variant) on the CSR pattern; columns are sorted within rows.
        lu.assign(v, v + rp[n]);
        std::vector<int64_t> pos(n, -1);
        for (int64_t i = 0; i < n; ++i) {
            for (int64_t q = rp[i]; q < rp[i + 1]; ++q) pos[ci[q]] = q;
            for (int64_t q = rp[i]; q < rp[i + 1] && ci[q] < i; ++q) {
                const int64_t k = ci[q];
                const double lik = lu[q] / lu[diag[k]];
                lu[q] = lik;
                for (int64_t qq = diag[k] + 1; qq < rp[k + 1]; ++qq) {
                    const int64_t p = pos[ci[qq]];
                    if (p >= 0) lu[p] -= lik * lu[qq];
                }
            }
            for (int64_t q = rp[i]; q < rp[i + 1]; ++q) pos[ci[q]] = -1;
        }
    }
    void apply(const double* x, double* y) const {
        if (kind == 0) {
#pragma omp parallel for schedule(static)
            for (int64_t r = 0; r < n; ++r) y[r] = dinv[r] * x[r];
            return;
        }
        for (int64_t i = 0; i < n; ++i) {
            double s = x[i];
            for (int64_t q = rp[i]; q < diag[i]; ++q) s -= lu[q] * y[ci[q]];
            y[i] = s;
        }
        for (int64_t i = n - 1; i >= 0; --i) {
            double s = y[i];
            for (int64_t q = diag[i] + 1; q < rp[i + 1]; ++q) s -= lu[q] * y[ci[q]];
            y[i] = s / lu[diag[i]];
        }
    }
};

// Right-preconditioned BiCGSTAB (van der Vorst 1992).  Stops on ||b - A x|| <= rtol ||b||,
// verified on the true residual.
static int bicgstab(int64_t n, const int* rp, const int* ci, const double* v, const double* b, double* x,
                    int precKind, double rtol, int maxit, int* itersOut, double* relresOut) {
    Precond M;
    M.setup(precKind, n, rp, ci, v);
    std::vector<double> r(n), rh(n), p(n, 0), vv(n, 0), ph(n), s(n), sh(n), t(n);
    const double bn = std::sqrt(dot(n, b, b));
    if (bn == 0) {
        std::fill(x, x + n, 0.0);
        *itersOut = 0;
        *relresOut = 0;
        return 0;
    }
    int it = 0;
    for (int restart = 0; restart < 20; ++restart) {
        spmv(n, rp, ci, v, x, r.data());
        for (int64_t i = 0; i < n; ++i) r[i] = b[i] - r[i];
        double rn = std::sqrt(dot(n, r.data(), r.data()));
        if (rn <= rtol * bn) {
            *itersOut = it;
            *relresOut = rn / bn;
            return 0;
        }
        if (it >= maxit) break;
        rh = r;
        double rho = 1, alpha = 1, omega = 1;
        std::fill(p.begin(), p.end(), 0.0);
        std::fill(vv.begin(), vv.end(), 0.0);
        bool breakdown = false;
        while (it < maxit) {
            double rhoNew = dot(n, rh.data(), r.data());
            if (rhoNew == 0 || !std::isfinite(rhoNew)) {
                breakdown = true;
                break;
            }
            const double beta = (rhoNew / rho) * (alpha / omega);
            rho = rhoNew;
#pragma omp parallel for schedule(static)
            for (int64_t i = 0; i < n; ++i) p[i] = r[i] + beta * (p[i] - omega * vv[i]);
            M.apply(p.data(), ph.data());
            spmv(n, rp, ci, v, ph.data(), vv.data());
            const double rv = dot(n, rh.data(), vv.data());
            alpha = rho / rv;
#pragma omp parallel for schedule(static)
            for (int64_t i = 0; i < n; ++i) s[i] = r[i] - alpha * vv[i];
            ++it;
            const double sn = std::sqrt(dot(n, s.data(), s.data()));
            if (sn <= rtol * bn) {
#pragma omp parallel for schedule(static)
                for (int64_t i = 0; i < n; ++i) x[i] += alpha * ph[i];
                break;
            }
            M.apply(s.data(), sh.data());
            spmv(n, rp, ci, v, sh.data(), t.data());
            const double tt = dot(n, t.data(), t.data());
            omega = (tt == 0) ? 0 : dot(n, t.data(), s.data()) / tt;
#pragma omp parallel for schedule(static)
            for (int64_t i = 0; i < n; ++i) {
                x[i] += alpha * ph[i] + omega * sh[i];
                r[i] = s[i] - omega * t[i];
            }
            const double rn2 = std::sqrt(dot(n, r.data(), r.data()));
            if (!std::isfinite(rn2)) return 2;
            if (rn2 <= rtol * bn || omega == 0) break;
        }
        (void)breakdown;
    }
    spmv(n, rp, ci, v, x, r.data());
    for (int64_t i = 0; i < n; ++i) r[i] = b[i] - r[i];
    const double rn = std::sqrt(dot(n, r.data(), r.data()));
    *itersOut = it;
    *relresOut = rn / bn;
    if (!std::isfinite(rn)) return 2;
    return rn <= rtol * bn ? 0 : 1;
}

// Dense LU with partial pivoting (tiny systems only).
static int denseSolve(int64_t n, const int* rp, const int* ci, const double* v, const double* b, double* x) {
    std::vector<double> A(n * n, 0.0);
    for (int64_t r = 0; r < n; ++r)
        for (int64_t k = rp[r]; k < rp[r + 1]; ++k) A[r * n + ci[k]] += v[k];
    std::vector<double> y(b, b + n);
    for (int64_t c = 0; c < n; ++c) {
        int64_t piv = c;
        for (int64_t r = c + 1; r < n; ++r)
            if (std::fabs(A[r * n + c]) > std::fabs(A[piv * n + c])) piv = r;
        if (A[piv * n + c] == 0) return 1;
        if (piv != c) {
            for (int64_t k = 0; k < n; ++k) std::swap(A[c * n + k], A[piv * n + k]);
            std::swap(y[c], y[piv]);
        }
        for (int64_t r = c + 1; r < n; ++r) {
            const double f = A[r * n + c] / A[c * n + c];
            if (f == 0) continue;
            for (int64_t k = c; k < n; ++k) A[r * n + k] -= f * A[c * n + k];
            y[r] -= f * y[c];
        }
    }
    for (int64_t r = n - 1; r >= 0; --r) {
        double s = y[r];
        for (int64_t k = r + 1; k < n; ++k) s -= A[r * n + k] * x[k];
        x[r] = s / A[r * n + r];
    }
    return 0;
}

struct SolveCfg {
    int kind;  // 0 dense LU, 1 BiCGSTAB+Jacobi, 2 BiCGSTAB+ILU(0)
    double rtol;
    int maxit;
};

static int linearSolve(const CsrSys& S, const SolveCfg& cfg, double* x, int* iters, double* relres) {
    if (cfg.kind == 0) {
        *iters = 0;
        int st = denseSolve(S.n, S.rowPtr.data(), S.colIdx.data(), S.vals.data(), S.b.data(), x);
        std::vector<double> r(S.n);
        spmv(S.n, S.rowPtr.data(), S.colIdx.data(), S.vals.data(), x, r.data());
        double rn = 0, bn = 0;
        for (int64_t i = 0; i < S.n; ++i) {
            rn += (S.b[i] - r[i]) * (S.b[i] - r[i]);
            bn += S.b[i] * S.b[i];
        }
        *relres = bn > 0 ? std::sqrt(rn / bn) : 0;
        return st;
    }
    return bicgstab(S.n, S.rowPtr.data(), S.colIdx.data(), S.vals.data(), S.b.data(), x, cfg.kind == 2 ? 1 : 0,
                    cfg.rtol, cfg.maxit, iters, relres);
}

template <int D>
static void assembleAndBC(const Params& P, const std::vector<uint8_t>& fl, const double* pos, const double* qPrev,
                          const double* qIter, const double* dirVals, CsrSys& S) {
    buildAbPSPG<D>(P, fl, pos, qPrev, qIter, S);
    applyBCPSPG<D>(P, fl, pos, qPrev, dirVals, S);
}

// Residual callback of the PSPG Picard loop (MomContEquationPSPG.inl:314-372).  kind 2
// ("Ax_f"): ||A q - b||_2 with the system re-assembled at the updated positions.  kind 0 ("U"):
// sqrt(sum |u - u_prev|^2 / sum |u_prev|^2) over the velocity components of the nodes that are
// not free, DBL_MAX when the denominator is 0; kind 1 ("U_P"): the larger of that and the same
// ratio over the pressures.
static double picardResidual(int D, int64_t N, const uint8_t* fl, int resKind, const CsrSys& S, const double* qIter,
                             const double* qIterPrev) {
    const int64_t n = (D + 1) * N;
    if (resKind == 2) {
        std::vector<double> Ax(n);
        spmv(n, S.rowPtr.data(), S.colIdx.data(), S.vals.data(), qIter, Ax.data());
        double s2 = 0;
        for (int64_t i = 0; i < n; ++i) s2 += (Ax[i] - S.b[i]) * (Ax[i] - S.b[i]);
        return std::sqrt(s2);
    }
    double num = 0, den = 0, resV, resP = 0;
    for (int64_t nn = 0; nn < N; ++nn)
        if (!(fl[nn] & FREE))
            for (int d = 0; d < D; ++d) {
                const double a = qIter[nn + d * N], b = qIterPrev[nn + d * N];
                num += (a - b) * (a - b);
                den += b * b;
            }
    resV = (den == 0) ? std::numeric_limits<double>::max() : std::sqrt(num / den);
    if (resKind == 1) {
        num = den = 0;
        for (int64_t nn = 0; nn < N; ++nn)
            if (!(fl[nn] & FREE)) {
                const double a = qIter[nn + D * N], b = qIterPrev[nn + D * N];
                num += (a - b) * (a - b);
                den += b * b;
            }
        resP = (den == 0) ? std::numeric_limits<double>::max() : std::sqrt(num / den);
    }
    return std::max(resV, resP);
}

// One PSPG time step: PicardAlgo::solve (srcs/simulation/nonLinearAlgo/PicardAlgo.cpp:31-95)
// with the PSPG callbacks (PSPG.inl:262-372).  On failure pos and q are left unchanged
// (restoreNodesList restores positions and states, Mesh.cpp:1019-1043).
struct StepStats {
    int picardIters = 0;
    int krylovIters = 0;
    int assemblies = 0;
    double finalRes = 0;
    double maxRelres = 0;
    double tAssemble = 0, tSolve = 0, tUpdate = 0, tResidual = 0;
    double tTriplets = 0, tPush = 0, tSetFromTriplets = 0, tVector = 0, tBC = 0;
};

// dirAffine (nullable): the Dirichlet values are re-evaluated before every assembly at the
// nodes' current positions, v_bc(x) = A x + c with A = dirAffine[d][0..D-1], c = dirAffine[d][D]
// per component d: the Lua "<Tag>V"(pos, t + dt) call of each m_applyBCPSPG
// (MomContEquationPSPG.inl:206-214) for a BC that depends on the position.
template <int D>
static int pspgStep(const Params& P, double* pos, double* q, const double* dirVals, int resKind, double minRes,
                    int maxIter, const SolveCfg& cfg, StepStats& st, const double* dirAffine = nullptr) {
    const int64_t N = P.N;
    const int64_t n = (D + 1) * N;
    auto fl = derivedFlags(P);
    std::vector<double> posSave(pos, pos + D * N);
    std::vector<double> qPrev(q, q + n);
    std::vector<double> qIter(n, 0.0), qIterPrev(qPrev);
    std::vector<double> states(qPrev);  // node states (u, v, [w], p) used by tau
    std::vector<double> dirEval;
    if (dirAffine) {
        dirEval.assign(dirVals, dirVals + D * N);
        dirVals = dirEval.data();
    }
    auto evalBC = [&]() {
        if (!dirAffine) return;
        for (int64_t nn = 0; nn < N; ++nn)
            if ((fl[nn] & BOUND) && (fl[nn] & DIRICHLET))
                for (int d = 0; d < D; ++d) {
                    double v = dirAffine[d * (D + 1) + D];
                    for (int k = 0; k < D; ++k) v += dirAffine[d * (D + 1) + k] * pos[nn + k * N];
                    dirEval[nn + d * N] = v;
                }
    };
    CsrSys S;
    auto restore = [&]() { std::copy(posSave.begin(), posSave.end(), pos); };
    auto t0 = clk::now();
    evalBC();
    assembleAndBC<D>(P, fl, pos, qPrev.data(), states.data(), dirVals, S);  // prepare
    st.assemblies++;
    st.tAssemble += secs(t0, clk::now());
    unsigned iterCount = 0;
    double res = std::numeric_limits<double>::max();
    std::vector<double> guess(qPrev);
    while (res > minRes) {
        if (iterCount > (unsigned)maxIter) {
            restore();
            st.picardIters = iterCount;
            st.finalRes = res;
            return 1;
        }
        // solve callback (PSPG.inl:278-313)
        auto ts = clk::now();
        std::vector<double> x(guess);
        int kit = 0;
        double rr = 0;
        int s = linearSolve(S, cfg, x.data(), &kit, &rr);
        st.krylovIters += kit;
        st.maxRelres = std::max(st.maxRelres, rr);
        st.tSolve += secs(ts, clk::now());
        if (s == 2 || (cfg.kind == 0 && s != 0)) {
            restore();
            st.picardIters = iterCount;
            return s == 2 ? 2 : 1;
        }
        qIter = x;
        guess = x;
        auto tu = clk::now();
        states = qIter;
        for (int64_t nn = 0; nn < N; ++nn)  // updateNodesPositionFromSave (Mesh.cpp:1238-1277)
            if (!(P.flags[nn] & FIXED))
                for (int d = 0; d < D; ++d) pos[nn + d * N] = posSave[nn + d * N] + qIter[nn + d * N] * P.dt;
        st.tUpdate += secs(tu, clk::now());
        auto ta = clk::now();
        S = CsrSys();
        evalBC();
        assembleAndBC<D>(P, fl, pos, qPrev.data(), states.data(), dirVals, S);
        st.assemblies++;
        st.tAssemble += secs(ta, clk::now());
        st.tTriplets += S.tTriplets;
        st.tPush += S.tPush;
        st.tSetFromTriplets += S.tAssemble;
        st.tVector += S.tVector;
        st.tBC += S.tBC;
        // residual callback (PSPG.inl:314-372)
        auto tr = clk::now();
        res = picardResidual(D, N, fl.data(), resKind, S, qIter.data(), qIterPrev.data());
        st.tResidual += secs(tr, clk::now());
        qIterPrev = qIter;
        if (std::isnan(res)) {
            restore();
            st.picardIters = iterCount + 1;
            st.finalRes = res;
            return 2;
        }
        iterCount++;
    }
    std::copy(qIter.begin(), qIter.end(), q);
    st.picardIters = iterCount;
    st.finalRes = res;
    return 0;
}

// --------------------------------------------------------------------------------------
// Weakly compressible CDS_dpdt step: srcs/simulation/physics/WCompNewton/Solver.cpp:236-276,
// ContEquation.inl:123-166,334-413 (Meduri / None stabilisation, Tait-Murnaghan EOS
// :304-332), MomEquation.inl:201-374.
// --------------------------------------------------------------------------------------
struct WCParams {
    int dim;
    int64_t N, E;
    const int* conn;
    const uint8_t* flags;
    double mu, g[3], K0, K0p, rhoStar;
    int meduri;
};

// Continuity system of one CDS_dpdt step (m_buildSystemdpdt, ContEquation.inl:353-413), scattered
// in element order: ML = the lump2 row sums of the consistent mass (factor 1,
// MatricesBuilder.hpp:87-102) and F0 = sum_e (-dt D(K0 + K0' N.P) V + M P) (Meduri) or
// (-dt D V + M_L P) (None).  Before m_applyBCdpdt and the inversion of ML.
template <int D>
static void wcContAssemble(const WCParams& W, double dt, const double* pos, const double* v, const double* p,
                           double* MLout, double* F0out) {
    constexpr int ne = D + 1, nd = D * ne, nv = MB<D>::nv;
    const int64_t N = W.N, E = W.E;
    MB<D> mb;
    std::vector<double> MLe(E * ne), F0e(E * ne);
#pragma omp parallel for schedule(static)
    for (int64_t e = 0; e < E; ++e) {
        const int* nodes = W.conn + e * ne;
        Geo geo;
        computeGeo<D>(pos, N, nodes, geo);
        double Me[ne][ne];
        mb.M(geo.detJ, [](const double*) { return 1.0; }, Me);
        double ML[ne];
        for (int i = 0; i < ne; ++i) {  // lump2: MatricesBuilder.hpp:87-102
            ML[i] = 0;
            for (int j = 0; j < ne; ++j) ML[i] += Me[i][j];
        }
        double Pe[ne], Ve[nd];
        for (int i = 0; i < ne; ++i) Pe[i] = p[nodes[i]];
        for (int d = 0; d < D; ++d)
            for (int i = 0; i < ne; ++i) Ve[i + d * ne] = v[nodes[i] + d * N];
        double G[D][ne], Bm[nv][nd], De[ne][nd];
        mb.gradN(geo, G);
        mb.B(G, Bm);
        mb.Dm(geo.detJ,
              [&](const double* Ng) {  // factor K0 + K0p * (N . P): ContEquation.inl:80-86
                  double pg = 0;
                  for (int i = 0; i < ne; ++i) pg += Ng[i] * Pe[i];
                  return W.K0 + W.K0p * pg;
              },
              Bm, De);
        for (int i = 0; i < ne; ++i) {
            double dv = 0;
            for (int c = 0; c < nd; ++c) dv += De[i][c] * Ve[c];
            double mp = 0;
            if (W.meduri)
                for (int j = 0; j < ne; ++j) mp += Me[i][j] * Pe[j];
            else
                mp = ML[i] * Pe[i];
            F0e[e * ne + i] = -dt * dv + mp;
            MLe[e * ne + i] = ML[i];
        }
    }
    std::fill(MLout, MLout + N, 0.0);
    std::fill(F0out, F0out + N, 0.0);
    for (int64_t e = 0; e < E; ++e)
        for (int i = 0; i < ne; ++i) {
            MLout[W.conn[e * ne + i]] += MLe[e * ne + i];
            F0out[W.conn[e * ne + i]] += F0e[e * ne + i];
        }
}

// Momentum system of one step (m_buildSystem, MomEquation.inl:229-302), scattered in element
// order: Mv = diagBlock + lump of M(rho_gp) (MatricesBuilder.hpp:104-129) per component, and
// F = sum_e (-K V + D^T P + F(rho_gp)) (K with the deviatoric ddev, MomEquation.inl:63-83).
// Before m_applyBC and the inversion of Mv.
template <int D>
static void wcMomAssemble(const WCParams& W, const double* pos, const double* v, const double* p, const double* rho,
                          double* Mvout, double* Fout) {
    constexpr int ne = D + 1, nd = D * ne, nv = MB<D>::nv;
    const int64_t N = W.N, E = W.E;
    MB<D> mb;
    std::vector<double> Mle(E * nd), Fte(E * nd);
    double dd[nv][nv];
    ddevWC<D>(dd);
#pragma omp parallel for schedule(static)
    for (int64_t e = 0; e < E; ++e) {
        const int* nodes = W.conn + e * ne;
        Geo geo;
        computeGeo<D>(pos, N, nodes, geo);
        double Ve[nd], Pe[ne], Re[ne];
        for (int d = 0; d < D; ++d)
            for (int i = 0; i < ne; ++i) Ve[i + d * ne] = v[nodes[i] + d * N];
        for (int i = 0; i < ne; ++i) {
            Pe[i] = p[nodes[i]];
            Re[i] = rho[nodes[i]];
        }
        auto rhoG = [&](const double* Ng) {
            double r = 0;
            for (int i = 0; i < ne; ++i) r += Ng[i] * Re[i];
            return r;
        };
        double G[D][ne], Bm[nv][nd];
        mb.gradN(geo, G);
        mb.B(G, Bm);
        double Ms[ne][ne];
        mb.M(geo.detJ, rhoG, Ms);
        // diagBlock then lump (row sums onto the diagonal): MatricesBuilder.hpp:104-129
        double lumped[ne];
        for (int i = 0; i < ne; ++i) {
            double s = Ms[i][i];
            for (int j = 0; j < ne; ++j)
                if (j != i) s += Ms[i][j];
            lumped[i] = s;
        }
        double Ke[nd][nd], De[ne][nd], Fe[nd];
        mb.K(geo.detJ, W.mu, dd, Bm, Ke);
        mb.Dm(geo.detJ, [](const double*) { return 1.0; }, Bm, De);
        mb.Fv(geo.detJ, rhoG, W.g, Fe);
        for (int r = 0; r < nd; ++r) {
            double kv = 0;
            for (int c = 0; c < nd; ++c) kv += Ke[r][c] * Ve[c];
            double dp = 0;
            for (int j = 0; j < ne; ++j) dp += De[j][r] * Pe[j];
            Fte[e * nd + r] = -kv + dp + Fe[r];
            Mle[e * nd + r] = lumped[r % ne];
        }
    }
    std::fill(Mvout, Mvout + D * N, 0.0);
    std::fill(Fout, Fout + D * N, 0.0);
    for (int64_t e = 0; e < E; ++e)
        for (int i = 0; i < ne; ++i)
            for (int d = 0; d < D; ++d) {
                const int64_t gi = W.conn[e * ne + i] + d * N;
                Mvout[gi] += Mle[e * nd + i + d * ne];
                Fout[gi] += Fte[e * nd + i + d * ne];
            }
}

template <int D>
static void wcStep(const WCParams& W, double dt, double* pos, double* v, double* p, double* rho, double* a,
                   const double* dirVals, double* tCont, double* tMom) {
    const int64_t N = W.N, E = W.E;
    Params P{};
    P.dim = D;
    P.N = N;
    P.E = E;
    P.conn = W.conn;
    P.flags = W.flags;
    auto fl = derivedFlags(P);
    // v1/2 = v + dt/2 a ; x += dt v1/2 (non-fixed)   Solver.cpp:249-254, Mesh.cpp:1101-1137
    for (int64_t i = 0; i < D * N; ++i) v[i] = v[i] + 0.5 * dt * a[i];
    for (int64_t nn = 0; nn < N; ++nn)
        if (!(W.flags[nn] & FIXED))
            for (int d = 0; d < D; ++d) pos[nn + d * N] += v[nn + d * N] * dt;
    auto t0 = clk::now();
    // ---- continuity (ContEquation.inl:353-413, m_applyBCdpdt :334-350, solve :123-166)
    std::vector<double> invM(N), F0(N);
    wcContAssemble<D>(W, dt, pos, v, p, invM.data(), F0.data());
    for (int64_t nn = 0; nn < N; ++nn) invM[nn] = 1 / invM[nn];
    for (int64_t nn = 0; nn < N; ++nn)  // m_applyBCdpdt: ContEquation.inl:334-350
        if (fl[nn] & FREE) {
            F0[nn] = 0;
            invM[nn] = 1;
        }
    for (int64_t nn = 0; nn < N; ++nn) {
        p[nn] = invM[nn] * F0[nn];
        rho[nn] = std::pow((W.K0p / W.K0) * p[nn] + 1, 1 / W.K0p) * W.rhoStar;  // :319-332
    }
    auto t1 = clk::now();
    // ---- momentum (MomEquation.inl:229-302, m_applyBC :305-374, solve :201-226)
    std::vector<double> invMv(D * N), F(D * N);
    wcMomAssemble<D>(W, pos, v, p, rho, invMv.data(), F.data());
    for (int64_t i = 0; i < D * N; ++i) invMv[i] = 1 / invMv[i];
    {  // the FST facet loop of m_applyBC (MomEquation.inl:312-335): facets with every node on the surface
        Material wm;
        addFST<D>(pos, N, W.conn, fl, wm, true, F.data());
    }
    for (int64_t nn = 0; nn < N; ++nn) {  // m_applyBC: MomEquation.inl:336-374
        const uint8_t f = fl[nn];
        if ((f & FREE) && !(f & BOUND)) {
            for (int d = 0; d < D; ++d) {
                F[nn + d * N] = W.g[d];
                invMv[nn + d * N] = 1;
            }
        } else if ((f & BOUND) && (f & DIRICHLET)) {
            for (int d = 0; d < D; ++d) {
                F[nn + d * N] = dirVals[nn + d * N];
                invMv[nn + d * N] = 1;
            }
        }
    }
    for (int64_t i = 0; i < D * N; ++i) {  // MomEquation.inl:216-222
        a[i] = invMv[i] * F[i];
        v[i] = v[i] + 0.5 * dt * a[i];
    }
    auto t2 = clk::now();
    if (tCont) *tCont += secs(t0, t1);
    if (tMom) *tMom += secs(t1, t2);
}

// computeNextDT: WCompNewton/Solver.cpp:192-234 (without the heat equation).
template <int D>
static double wcNextDT(const WCParams& W, const double* pos, const double* v, const double* p, const double* rho,
                       double securityCoeff, double maxDT) {
    constexpr int ne = D + 1;
    const int64_t N = W.N;
    double timeStep = std::numeric_limits<double>::max();
#pragma omp parallel for reduction(min : timeStep)
    for (int64_t e = 0; e < W.E; ++e) {
        const int* nodes = W.conn + e * ne;
        Geo geo;
        computeGeo<D>(pos, N, nodes, geo);
        const double he = 2 * getRin<D>(pos, N, nodes, geo.detJ);
        double maxS = 0;
        for (int k = 0; k < ne; ++k) {
            const int nn = nodes[k];
            const double c2 = (W.K0 + W.K0p * p[nn]) / rho[nn];  // ContEquation.inl:117-121
            double u2 = 0;
            for (int d = 0; d < D; ++d) u2 += v[nn + d * N] * v[nn + d * N];  // MomEquation.inl:184-192
            const double alpha = W.mu / rho[nn];                              // MomEquation.inl:195-198
            maxS = std::max(std::max(u2, c2), maxS);
            maxS = std::max(maxS, 4 * alpha * alpha / (he * he));
        }
        timeStep = std::min(timeStep, securityCoeff * securityCoeff * he * he / maxS);
    }
    return std::min(std::sqrt(timeStep), maxDT);
}

}  // namespace

// ======================================================================================
// C interface (ctypes) — used by tests/, __graft_entry__.smoke() and bench.py's CPU leg.
// ======================================================================================
// --------------------------------------------------------------------------------------
// Heat equation (conduction, no phase change): srcs/simulation/physics/IncompNewton/
// HeatEquation.inl:6-425 — m_buildAb (:209-280), m_applyBC (:283-364), the Picard callbacks
// with runOnlyOnce (:111-194) and Eigen::ConjugateGradient<Lower|Upper> (HeatEquation.hpp:50)
// with its default diagonal preconditioner, solveWithGuess(b, qPrev) (:131).
// --------------------------------------------------------------------------------------
enum HeatFacetKind : uint8_t { HEAT_Q = 1, HEAT_QH = 2, HEAT_QR = 4 };

struct HeatParams {
    int dim = 2;
    int64_t N = 0, E = 0, F = 0;
    const int* conn = nullptr;
    const uint8_t* flags = nullptr;  // DIRICHLET: the node's tag has a "<Tag>T" function
    double rho = 0, cv = 0, k = 0, h = 0, Tinf = 0, epsRad = 0, dt = 0;
    const int* facets = nullptr;     // F x dim
    const int* fout = nullptr;       // F: the facet's out node (Facet::m_outNodeIndex)
    const uint8_t* fkind = nullptr;  // F: HEAT_* bits (the facet's nodes all carry that BC)
    const double* qgp = nullptr;     // F x 3 x dim: "<Tag>Q"(gauss point, t + dt), nullable
};

// facet quadrature, nGPLD = 3 (HeatEquation.inl:12-23): Mesh.cpp:342-530 for dimension dim-1
// m_buildAb: HeatEquation.inl:209-280 (triplets M then dt L per node pair of rows that are
// neither Dirichlet nor free; unit diagonals pushed for Dirichlet or free nodes; b = sum of
// M theta_prev per element; setFromTriplets as in buildAbPSPG, the default-triplet padding
// keeping (0, 0) structural)
template <int D>
static void buildAbHeat(const HeatParams& P, const std::vector<uint8_t>& fl, const double* pos, const double* qPrev,
                        CsrSys& S) {
    constexpr int ne = D + 1;
    const int64_t N = P.N, E = P.E;
    MB<D> mb;
    std::vector<Triplet> indexA((size_t)E * 2 * ne * ne + 1, Triplet{0, 0, 0.0});
    std::vector<std::pair<int64_t, double>> indexb((size_t)E * ne);
#pragma omp parallel for schedule(static)
    for (int64_t elm = 0; elm < E; ++elm) {
        const int* nodes = P.conn + elm * ne;
        Geo geo;
        computeGeo<D>(pos, N, nodes, geo);
        double G[D][ne], Me[ne][ne], Le[ne][ne];
        mb.gradN(geo, G);
        mb.M(geo.detJ, [&](const double*) { return P.cv * P.rho; }, Me);
        mb.L(geo.detJ, P.k, G, Le);
        for (int i = 0; i < ne; ++i)
            for (int j = 0; j < ne; ++j) Le[i][j] = P.dt * Le[i][j];
        double th[ne], Mth[ne];
        for (int a = 0; a < ne; ++a) th[a] = qPrev[nodes[a]];
        for (int i = 0; i < ne; ++i) {
            double t = 0;
            for (int j = 0; j < ne; ++j) t += Me[i][j] * th[j];
            Mth[i] = t;
        }
        int64_t countA = 0;
        for (int i = 0; i < ne; ++i) {
            const uint8_t fi = fl[nodes[i]];
            for (int j = 0; j < ne; ++j) {
                if (!(fi & DIRICHLET)) {
                    if (!(fi & FREE))
                        indexA[(size_t)2 * ne * ne * elm + countA] = Triplet{nodes[i], nodes[j], Me[i][j]};
                    countA++;
                    if (!(fi & FREE))
                        indexA[(size_t)2 * ne * ne * elm + countA] = Triplet{nodes[i], nodes[j], Le[i][j]};
                    countA++;
                }
            }
            indexb[(size_t)ne * elm + i] = {nodes[i], Mth[i]};
        }
    }
    for (int64_t nn = 0; nn < N; ++nn)
        if ((fl[nn] & DIRICHLET) || (fl[nn] & FREE)) indexA.push_back(Triplet{nn, nn, 1});
    std::vector<int64_t> cnt(N + 1, 0);
    for (const auto& t : indexA) cnt[t.row + 1]++;
    for (int64_t r = 0; r < N; ++r) cnt[r + 1] += cnt[r];
    std::vector<int64_t> cur(cnt.begin(), cnt.end() - 1);
    std::vector<int64_t> ecol(indexA.size());
    std::vector<double> eval(indexA.size());
    for (const auto& t : indexA) {
        const int64_t q = cur[t.row]++;
        ecol[q] = t.col;
        eval[q] = t.val;
    }
    S.n = N;
    S.rowPtr.assign(N + 1, 0);
    std::vector<std::vector<std::pair<int64_t, double>>> rows(N);
    for (int64_t r = 0; r < N; ++r) {
        auto& out = rows[r];
        for (int64_t q = cnt[r]; q < cnt[r + 1]; ++q) {
            bool found = false;
            for (auto& pr : out)
                if (pr.first == ecol[q]) {
                    pr.second += eval[q];
                    found = true;
                    break;
                }
            if (!found) out.push_back({ecol[q], eval[q]});
        }
        std::sort(out.begin(), out.end());
        S.rowPtr[r + 1] = S.rowPtr[r] + (int)out.size();
    }
    S.colIdx.resize(S.rowPtr[N]);
    S.vals.resize(S.rowPtr[N]);
    for (int64_t r = 0; r < N; ++r) {
        int64_t q = S.rowPtr[r];
        for (const auto& pr : rows[r]) {
            S.colIdx[q] = (int)pr.first;
            S.vals[q] = pr.second;
            ++q;
        }
    }
    S.b.assign(N, 0.0);
    for (const auto& db : indexb) S.b[db.first] += db.second;
}

// getQN (MatricesBuilder.inl:247-274) and getSGamma (:358-371) of one facet with the two
// SGamma factors of HeatEquation.inl:87-99 (convection h (T - Tinf), radiation
// sigma epsRad (T^4 - Tinf^4), T = N theta at each Gauss point)
template <int D>
static void heatFacetTerms(const HeatParams& P, const double* pos, const double* T, int64_t f, double qn[D],
                           double sgh[D], double sgr[D]) {
    QuadLD<D> q;
    const int* fn = P.facets + f * D;
    double detJ, nrm[3];
    facetGeo<D>(pos, P.N, fn, P.fout[f], detJ, nrm);
    const double s = q.refSize * detJ;
    for (int i = 0; i < D; ++i) qn[i] = sgh[i] = sgr[i] = 0;
    for (int g = 0; g < 3; ++g) {
        double qv = 0;
        if (P.qgp)
            for (int k = 0; k < D; ++k) qv += P.qgp[(f * 3 + g) * D + k] * nrm[k];
        const double fact = qv * q.w[g];
        double Tg = 0;
        for (int j = 0; j < D; ++j) Tg += q.N[g][j] * T[fn[j]];
        const double fh = P.h * (Tg - P.Tinf);
        const double sigma = 5.670374419e-8;
        const double fr = sigma * P.epsRad * (Tg * Tg * Tg * Tg - P.Tinf * P.Tinf * P.Tinf * P.Tinf);
        for (int i = 0; i < D; ++i) {
            qn[i] += fact * q.N[g][i];
            sgh[i] += fh * q.N[g][i] * q.w[g];
            sgr[i] += fr * q.N[g][i] * q.w[g];
        }
    }
    for (int i = 0; i < D; ++i) {
        qn[i] = qn[i] * s;
        sgh[i] *= s;
        sgr[i] *= s;
    }
}

// m_applyBC: HeatEquation.inl:283-364 (facet fluxes in facet order, then free-node rhs and the
// Dirichlet column elimination in node order)
template <int D>
static void applyBCHeat(const HeatParams& P, const std::vector<uint8_t>& fl, const double* pos, const double* qPrev,
                        const double* Tbc, CsrSys& S) {
    for (int64_t f = 0; f < P.F; ++f) {
        const uint8_t kind = P.fkind[f];
        if (!kind) continue;
        double qn[D], sgh[D], sgr[D];
        heatFacetTerms<D>(P, pos, qPrev, f, qn, sgh, sgr);
        const int* fn = P.facets + f * D;
        if (kind & HEAT_Q)
            for (int i = 0; i < D; ++i) S.b[fn[i]] -= P.dt * qn[i];
        if (kind & HEAT_QH)
            for (int i = 0; i < D; ++i) S.b[fn[i]] -= P.dt * sgh[i];
        if (kind & HEAT_QR)
            for (int i = 0; i < D; ++i) S.b[fn[i]] -= P.dt * sgr[i];
    }
    const int64_t N = P.N;
    std::vector<int64_t> ccnt(N + 1, 0);
    for (int c : S.colIdx) ccnt[c + 1]++;
    for (int64_t c = 0; c < N; ++c) ccnt[c + 1] += ccnt[c];
    std::vector<int64_t> cpos(ccnt.begin(), ccnt.end() - 1), cRow(S.colIdx.size()), cK(S.colIdx.size());
    for (int64_t r = 0; r < N; ++r)
        for (int64_t q = S.rowPtr[r]; q < S.rowPtr[r + 1]; ++q) {
            const int64_t c = S.colIdx[q];
            cRow[cpos[c]] = r;
            cK[cpos[c]] = q;
            cpos[c]++;
        }
    for (int64_t nn = 0; nn < N; ++nn) {
        if ((fl[nn] & FREE) && !(fl[nn] & DIRICHLET)) {
            S.b[nn] = qPrev[nn];
        } else if (fl[nn] & DIRICHLET) {
            const double res = Tbc[nn];
            S.b[nn] = res;
            for (int64_t q = ccnt[nn]; q < ccnt[nn + 1]; ++q) {
                const int64_t row = cRow[q];
                if (row == nn) continue;
                S.b[row] -= S.vals[cK[q]] * res;
                S.vals[cK[q]] = 0;
            }
        }
    }
}

// Conjugate gradient with the diagonal preconditioner (Eigen's ConjugateGradient default), from
// the initial guess x; stops on ||b - A x|| <= rtol ||b|| or after maxit iterations.
static int pcg(int64_t n, const int* rp, const int* ci, const double* v, const double* b, double* x, double rtol,
               int maxit, int* itersOut, double* relresOut) {
    std::vector<double> dinv(n, 1.0), r(n), z(n), p(n), q(n);
    for (int64_t i = 0; i < n; ++i)
        for (int64_t k = rp[i]; k < rp[i + 1]; ++k)
            if (ci[k] == i && v[k] != 0) dinv[i] = 1.0 / v[k];
    spmv(n, rp, ci, v, x, q.data());
    for (int64_t i = 0; i < n; ++i) r[i] = b[i] - q[i];
    const double bb = dot(n, b, b);
    if (bb == 0) {
        std::fill(x, x + n, 0.0);
        *itersOut = 0;
        *relresOut = 0;
        return 0;
    }
    const double thr = rtol * rtol * bb;
    double rr = dot(n, r.data(), r.data());
    int it = 0;
    if (rr >= thr) {
        for (int64_t i = 0; i < n; ++i) p[i] = z[i] = dinv[i] * r[i];
        double rz = dot(n, r.data(), z.data());
        for (; it < maxit;) {
            spmv(n, rp, ci, v, p.data(), q.data());
            const double alpha = rz / dot(n, p.data(), q.data());
            for (int64_t i = 0; i < n; ++i) {
                x[i] += alpha * p[i];
                r[i] -= alpha * q[i];
            }
            ++it;
            rr = dot(n, r.data(), r.data());
            if (rr < thr) break;
            for (int64_t i = 0; i < n; ++i) z[i] = dinv[i] * r[i];
            const double rzNew = dot(n, r.data(), z.data());
            const double beta = rzNew / rz;
            rz = rzNew;
            for (int64_t i = 0; i < n; ++i) p[i] = z[i] + beta * p[i];
        }
    }
    *itersOut = it;
    *relresOut = std::sqrt(rr / bb);
    return std::isfinite(rr) ? (rr < thr ? 0 : 1) : 2;
}

template <int D>
static void heatSystem(const HeatParams& P, const double* pos, const double* qPrev, const double* Tbc, CsrSys& S) {
    std::vector<uint8_t> fl(P.flags, P.flags + P.N);
    std::vector<char> used(P.N, 0);
    for (int64_t q = 0; q < P.E * (D + 1); ++q) used[P.conn[q]] = 1;
    for (int64_t n = 0; n < P.N; ++n) fl[n] = (fl[n] & ~FREE) | (used[n] ? 0 : FREE);
    buildAbHeat<D>(P, fl, pos, qPrev, S);
    applyBCHeat<D>(P, fl, pos, qPrev, Tbc, S);
}

// --------------------------------------------------------------------------------------
// Fractional step (Solver.id = "FracStep"): srcs/simulation/physics/IncompNewton/
// MomContEquationFracStep.inl — m_buildMatFracStep (:7-195), m_applyBCVAppStep (:197-266, the
// FST loop skipped at gamma = 0), m_buildMatPcorrStep (:268-314), m_applyBCPCorrStep (:316-341),
// m_buildMatVStep (:343-429), m_applyBCVStep (:431-460), the Picard callbacks (:462-630): three
// ConjugateGradient<Lower|Upper> solves per iteration (MomContEquation.hpp:49) from a zero guess.
// --------------------------------------------------------------------------------------
// setFromTriplets as in buildAbPSPG (duplicates summed in triplet order, columns sorted)
static void csrFromTriplets(int64_t n, const std::vector<Triplet>& T, CsrSys& S) {
    std::vector<int64_t> cnt(n + 1, 0);
    for (const auto& t : T) cnt[t.row + 1]++;
    for (int64_t r = 0; r < n; ++r) cnt[r + 1] += cnt[r];
    std::vector<int64_t> cur(cnt.begin(), cnt.end() - 1), ecol(T.size());
    std::vector<double> eval(T.size());
    for (const auto& t : T) {
        const int64_t q = cur[t.row]++;
        ecol[q] = t.col;
        eval[q] = t.val;
    }
    S.n = n;
    S.rowPtr.assign(n + 1, 0);
    std::vector<std::vector<std::pair<int64_t, double>>> rows(n);
#pragma omp parallel for schedule(dynamic, 1024)
    for (int64_t r = 0; r < n; ++r) {
        auto& out = rows[r];
        for (int64_t q = cnt[r]; q < cnt[r + 1]; ++q) {
            bool found = false;
            for (auto& pr : out)
                if (pr.first == ecol[q]) {
                    pr.second += eval[q];
                    found = true;
                    break;
                }
            if (!found) out.push_back({ecol[q], eval[q]});
        }
        std::sort(out.begin(), out.end());
    }
    for (int64_t r = 0; r < n; ++r) S.rowPtr[r + 1] = S.rowPtr[r] + (int)rows[r].size();
    S.colIdx.resize(S.rowPtr[n]);
    S.vals.resize(S.rowPtr[n]);
    for (int64_t r = 0; r < n; ++r) {
        int64_t q = S.rowPtr[r];
        for (const auto& pr : rows[r]) {
            S.colIdx[q] = (int)pr.first;
            S.vals[q] = pr.second;
            ++q;
        }
    }
}

struct FracSys {
    CsrSys M, MK, L;                   // m_M, m_MK_dt (with bVAppStep in MK.b), m_L
    std::vector<std::array<double, 12 * 4>> DT;  // m_DTelm: D^T (d ne x ne) per element, row-major
    std::vector<std::array<double, 16>> Le;      // m_Lelm (ne x ne)
};

// m_buildMatFracStep + m_applyBCVAppStep
template <int D>
static void buildFracVApp(const Params& P, const std::vector<uint8_t>& fl, const double* pos, const double* vPrevQ,
                          const double* pPrevQ, const double* statesV, const double* dirVals, double gammaFS,
                          FracSys& F) {
    constexpr int ne = D + 1, nd = D * ne, nv = MB<D>::nv;
    const int64_t N = P.N, E = P.E;
    const double dt = P.dt;
    MB<D> mb;
    double dd[nv][nv];
    ddevIncomp<D>(dd);
    const int64_t tpM = (int64_t)D * ne * ne, tpMK = (int64_t)D * ne * ne + (int64_t)D * ne * D * ne,
                  tpL = (int64_t)ne * ne, dpb = 3 * (int64_t)D * ne;
    std::vector<Triplet> iM(tpM * E), iMK(tpMK * E), iL(tpL * E);
    std::vector<std::pair<int64_t, double>> ib(dpb * E, {0, 0.0});
    F.DT.assign(E, {});
    F.Le.assign(E, {});
#pragma omp parallel for schedule(static)
    for (int64_t elm = 0; elm < E; ++elm) {
        const int* nodes = P.conn + elm * ne;
        Geo geo;
        computeGeo<D>(pos, N, nodes, geo);
        double G[D][ne], Bm[nv][nd];
        mb.gradN(geo, G);
        mb.B(G, Bm);
        double Ms[ne][ne], Ke[nd][nd], Dm[ne][nd], Le[ne][ne], Fe[nd];
        mb.M(geo.detJ, [&](const double*) { return P.rho; }, Ms);
        mb.K(geo.detJ, kFactor<D>(P, Bm, dd, nodes, statesV), dd, Bm, Ke);
        mb.Dm(geo.detJ, [&](const double*) { return 1.0; }, Bm, Dm);
        mb.L(geo.detJ, 1.0, G, Le);
        mb.Fv(geo.detJ, [&](const double*) { return P.rho; }, P.g, Fe);
        const double idt = 1 / dt;
        double Ms_dt[ne][ne];
        for (int a = 0; a < ne; ++a)
            for (int b = 0; b < ne; ++b) Ms_dt[a][b] = idt * Ms[a][b];
        for (int r = 0; r < nd; ++r)
            for (int c = 0; c < ne; ++c) F.DT[elm][r * ne + c] = Dm[c][r];
        for (int a = 0; a < ne; ++a)
            for (int b = 0; b < ne; ++b) F.Le[elm][a * ne + b] = Le[a][b];
        double vP[nd], pP[ne];
        for (int d = 0; d < D; ++d)
            for (int a = 0; a < ne; ++a) vP[a + d * ne] = vPrevQ[nodes[a] + d * N];
        for (int a = 0; a < ne; ++a) pP[a] = pPrevQ[nodes[a]];
        double Mv[nd], DTp[nd];
        for (int d = 0; d < D; ++d)
            for (int a = 0; a < ne; ++a) {
                double t = 0;
                for (int b = 0; b < ne; ++b) t += Ms_dt[a][b] * vP[b + d * ne];
                Mv[a + d * ne] = t;
            }
        for (int r = 0; r < nd; ++r) {
            double t = 0;
            for (int c = 0; c < ne; ++c) t += F.DT[elm][r * ne + c] * pP[c];
            DTp[r] = gammaFS * t;
        }
        int64_t cM = 0, cMK = 0, cL = 0, cb = 0;
        for (int i = 0; i < ne; ++i) {
            const uint8_t fi = fl[nodes[i]];
            const bool bf = (fi & (BOUND | FREE)) != 0;
            for (int j = 0; j < ne; ++j) {
                for (int d = 0; d < D; ++d) {
                    if (!bf) {
                        iM[tpM * elm + cM] = Triplet{nodes[i] + d * N, nodes[j] + d * N, Ms[i][j]};
                        iMK[tpMK * elm + cMK] = Triplet{nodes[i] + d * N, nodes[j] + d * N, Ms_dt[i][j]};
                    }
                    cM++;
                    cMK++;
                    for (int d2 = 0; d2 < D; ++d2) {
                        if (!bf)
                            iMK[tpMK * elm + cMK] =
                                Triplet{nodes[i] + d * N, nodes[j] + d2 * N, Ke[i + d * ne][j + d2 * ne]};
                        cMK++;
                    }
                }
                if (!(fi & FREE)) iL[tpL * elm + cL] = Triplet{nodes[i], nodes[j], Le[i][j]};
                cL++;
            }
            for (int d = 0; d < D; ++d)
                if (!bf) {
                    ib[dpb * elm + cb++] = {nodes[i] + d * N, Fe[i + d * ne]};
                    ib[dpb * elm + cb++] = {nodes[i] + d * N, Mv[i + d * ne]};
                    ib[dpb * elm + cb++] = {nodes[i] + d * N, DTp[i + d * ne]};
                }
        }
    }
    for (int64_t nn = 0; nn < N; ++nn) {
        if (fl[nn] & FREE) iL.push_back(Triplet{nn, nn, 1});
        if (fl[nn] & (BOUND | FREE))
            for (int d = 0; d < D; ++d) {
                iMK.push_back(Triplet{nn + d * N, nn + d * N, 1});
                iM.push_back(Triplet{nn + d * N, nn + d * N, 1});
            }
    }
    csrFromTriplets(D * N, iM, F.M);
    csrFromTriplets(D * N, iMK, F.MK);
    csrFromTriplets(N, iL, F.L);
    F.MK.b.assign(D * N, 0.0);
    for (const auto& db : ib) F.MK.b[db.first] += db.second;
    // m_applyBCVAppStep without its FST facet loop (surface tension is not restated for this
    // solver: its pressure correction has no free-surface condition), then the free and
    // Dirichlet rows
    CsrSys& S = F.MK;
    std::vector<int64_t> ccnt(S.n + 1, 0);
    for (int c : S.colIdx) ccnt[c + 1]++;
    for (int64_t c = 0; c < S.n; ++c) ccnt[c + 1] += ccnt[c];
    std::vector<int64_t> cpos(ccnt.begin(), ccnt.end() - 1), cRow(S.colIdx.size()), cK(S.colIdx.size());
    for (int64_t r = 0; r < S.n; ++r)
        for (int64_t q = S.rowPtr[r]; q < S.rowPtr[r + 1]; ++q) {
            const int64_t c = S.colIdx[q];
            cRow[cpos[c]] = r;
            cK[cpos[c]] = q;
            cpos[c]++;
        }
    for (int64_t nn = 0; nn < N; ++nn) {
        if ((fl[nn] & FREE) && !(fl[nn] & BOUND))
            for (int d = 0; d < D; ++d) S.b[nn + d * N] = vPrevQ[nn + d * N] + dt * P.g[d];
        if ((fl[nn] & BOUND) && (fl[nn] & DIRICHLET))
            for (int d = 0; d < D; ++d) {
                const double res = dirVals[nn + d * N];
                const int64_t col = nn + d * N;
                S.b[col] = res;
                for (int64_t q = ccnt[col]; q < ccnt[col + 1]; ++q) {
                    const int64_t row = cRow[q];
                    if (row == col) continue;
                    S.b[row] -= S.vals[cK[q]] * res;
                    S.vals[cK[q]] = 0;
                }
            }
    }
}

// column c of a CSR system set to the unit column (diagonal 1, other entries 0)
static void unitColumns(CsrSys& S, const std::vector<char>& isUnit) {
    for (int64_t r = 0; r < S.n; ++r)
        for (int64_t q = S.rowPtr[r]; q < S.rowPtr[r + 1]; ++q) {
            const int64_t c = S.colIdx[q];
            if (isUnit[c]) S.vals[q] = (c == r) ? 1.0 : 0.0;
        }
}

struct FracStats {
    int picardIters = 0, cgIters[3] = {0, 0, 0};
    double finalRes = 0;
};

// One FracStep time step (PicardAlgo::solve with the callbacks of :462-630); q = (v, p)
// field-blocked (d+1) N, updated with pos on success.  Returns 0, 1 (Picard not converged) or 2
// (NaN).  Each solve is CG to rtol, at most cgMaxit iterations (2 n when cgMaxit <= 0, Eigen's
// default); like the reference, which checks only the preconditioner's set-up (:489-504), an
// unconverged solve is used as it stands (the pressure Laplacian is singular: a right-hand side
// off its range, as on a free surface, stalls CG).
template <int D>
static int fracStep(const Params& P, double* pos, double* q, const double* dirVals, double gammaFS, int resKind,
                    double minRes, int maxIter, double rtol, int cgMaxit, FracStats& st) {
    const int64_t N = P.N, nv = D * N;
    auto fl = derivedFlags(P);
    std::vector<double> posSave(pos, pos + D * N);
    std::vector<double> vPrev(q, q + nv), pPrev(q + nv, q + nv + N);
    std::vector<double> vIter(nv, 0.0), pIter(N, 0.0), vIterPrev(vPrev), pIterPrev(pPrev);
    std::vector<double> statesV(vPrev);  // node velocity states read by the K factor
    auto restore = [&]() { std::copy(posSave.begin(), posSave.end(), pos); };
    unsigned it = 0;
    double res = std::numeric_limits<double>::max();
    while (res > minRes) {
        if (it > (unsigned)maxIter) {
            restore();
            st.picardIters = it;
            return 1;
        }
        FracSys F;
        if (D == 2)
            buildFracVApp<2>(P, fl, pos, vPrev.data(), pPrev.data(), statesV.data(), dirVals, gammaFS, F);
        else
            buildFracVApp<3>(P, fl, pos, vPrev.data(), pPrev.data(), statesV.data(), dirVals, gammaFS, F);
        int its;
        double rr;
        std::vector<double> vTilde(nv, 0.0);
        if (pcg(nv, F.MK.rowPtr.data(), F.MK.colIdx.data(), F.MK.vals.data(), F.MK.b.data(), vTilde.data(), rtol,
                cgMaxit > 0 ? cgMaxit : (int)(2 * nv), &its, &rr) == 2) {
            restore();
            st.picardIters = it;
            return 2;
        }
        st.cgIters[0] += its;
        // m_buildMatPcorrStep + m_applyBCPCorrStep
        constexpr int ne = D + 1;
        std::vector<std::pair<int64_t, double>> ibp((size_t)2 * ne * P.E, {0, 0.0});
        for (int64_t elm = 0; elm < P.E; ++elm) {
            const int* nodes = P.conn + elm * ne;
            double vT[D * ne], pP[ne];
            for (int d = 0; d < D; ++d)
                for (int a = 0; a < ne; ++a) vT[a + d * ne] = vTilde[nodes[a] + d * N];
            for (int a = 0; a < ne; ++a) pP[a] = pPrev[nodes[a]];
            int64_t c = 0;
            for (int i = 0; i < ne; ++i) {
                double Dv = 0, Lp = 0;
                for (int r = 0; r < D * ne; ++r) Dv += F.DT[elm][r * ne + i] * vT[r];
                for (int b = 0; b < ne; ++b) Lp += F.Le[elm][i * ne + b] * pP[b];
                const uint8_t fi = fl[nodes[i]];
                if (!(fi & FREE) && !(fi & ONFS)) {
                    ibp[(size_t)2 * ne * elm + c++] = {nodes[i], -((P.rho / P.dt) * Dv)};
                    ibp[(size_t)2 * ne * elm + c++] = {nodes[i], gammaFS * Lp};
                }
            }
        }
        F.L.b.assign(N, 0.0);
        for (const auto& db : ibp) F.L.b[db.first] += db.second;
        std::vector<char> freeNode(N, 0);
        for (int64_t nn = 0; nn < N; ++nn)
            if (fl[nn] & FREE) {
                F.L.b[nn] = 0;
                freeNode[nn] = 1;
            }
        unitColumns(F.L, freeNode);
        std::vector<double> pNew(N, 0.0);
        if (pcg(N, F.L.rowPtr.data(), F.L.colIdx.data(), F.L.vals.data(), F.L.b.data(), pNew.data(), rtol,
                cgMaxit > 0 ? cgMaxit : (int)(2 * N), &its, &rr) == 2) {
            restore();
            st.picardIters = it;
            return 2;
        }
        st.cgIters[1] += its;
        // m_buildMatVStep + m_applyBCVStep
        std::vector<double> dP(N);
        for (int64_t nn = 0; nn < N; ++nn) dP[nn] = pNew[nn] - gammaFS * pPrev[nn];
        std::vector<std::pair<int64_t, double>> ibv((size_t)D * ne * P.E, {0, 0.0});
        for (int64_t elm = 0; elm < P.E; ++elm) {
            const int* nodes = P.conn + elm * ne;
            double dPe[ne];
            for (int a = 0; a < ne; ++a) dPe[a] = dP[nodes[a]];
            int64_t c = 0;
            for (int i = 0; i < ne; ++i) {
                const uint8_t fi = fl[nodes[i]];
                if (!(fi & FREE) && !(fi & BOUND))
                    for (int d = 0; d < D; ++d) {
                        double t = 0;
                        for (int b = 0; b < ne; ++b) t += F.DT[elm][(i + d * ne) * ne + b] * dPe[b];
                        ibv[(size_t)D * ne * elm + c++] = {nodes[i] + d * N, P.dt * t};
                    }
            }
        }
        F.M.b.assign(nv, 0.0);
        for (const auto& db : ibv) F.M.b[db.first] += db.second;
        std::vector<char> unitV(nv, 0);
        for (int64_t nn = 0; nn < N; ++nn)
            if (fl[nn] & (FREE | BOUND))
                for (int d = 0; d < D; ++d) {
                    F.M.b[nn + d * N] = 0;
                    unitV[nn + d * N] = 1;
                }
        unitColumns(F.M, unitV);
        std::vector<double> dV(nv, 0.0);
        if (pcg(nv, F.M.rowPtr.data(), F.M.colIdx.data(), F.M.vals.data(), F.M.b.data(), dV.data(), rtol,
                cgMaxit > 0 ? cgMaxit : (int)(2 * nv), &its, &rr) == 2) {
            restore();
            st.picardIters = it;
            return 2;
        }
        st.cgIters[2] += its;
        for (int64_t k = 0; k < nv; ++k) vIter[k] = vTilde[k] + dV[k];
        pIter = pNew;
        statesV = vIter;  // setNodesStatesfromQ (:558-559)
        for (int64_t nn = 0; nn < N; ++nn)  // updateNodesPositionFromSave (Mesh.cpp:1238-1277)
            if (!(P.flags[nn] & FIXED))
                for (int d = 0; d < D; ++d) pos[nn + d * N] = posSave[nn + d * N] + vIter[nn + d * N] * P.dt;
        // residual (:568-628): U or U_P over the nodes that are not free
        double num = 0, den = 0, resV, resP = 0;
        for (int64_t nn = 0; nn < N; ++nn)
            if (!(fl[nn] & FREE))
                for (int d = 0; d < D; ++d) {
                    const double a = vIter[nn + d * N], b = vIterPrev[nn + d * N];
                    num += (a - b) * (a - b);
                    den += b * b;
                }
        resV = den == 0 ? std::numeric_limits<double>::max() : std::sqrt(num / den);
        if (resKind == 1) {
            num = den = 0;
            for (int64_t nn = 0; nn < N; ++nn)
                if (!(fl[nn] & FREE)) {
                    num += (pIter[nn] - pIterPrev[nn]) * (pIter[nn] - pIterPrev[nn]);
                    den += pIterPrev[nn] * pIterPrev[nn];
                }
            resP = den == 0 ? std::numeric_limits<double>::max() : std::sqrt(num / den);
        }
        res = std::max(resV, resP);
        vIterPrev = vIter;
        pIterPrev = pIter;
        if (std::isnan(res)) {
            restore();
            st.picardIters = it + 1;
            return 2;
        }
        it++;
    }
    std::copy(vIter.begin(), vIter.end(), q);
    std::copy(pIter.begin(), pIter.end(), q + nv);
    st.picardIters = it;
    st.finalRes = res;
    return 0;
}

extern "C" {

struct orc_sys {
    CsrSys S;
};

static void setThreads(int nthreads) {
    if (nthreads > 0) omp_set_num_threads(nthreads);
}

// m_buildAbPSPG (+ m_applyBCPSPG when apply_bc != 0).
orc_sys* orc_pspg_assemble(int dim, int64_t N, int64_t E, const int* conn, const uint8_t* flags, const double* pos,
                           const double* q_prev, const double* q_iter, const double* dir_vals, double rho, double mu,
                           const double* g, double dt, int apply_bc, int nthreads) {
    setThreads(nthreads);
    Params P{};
    P.dim = dim;
    P.N = N;
    P.E = E;
    P.conn = conn;
    P.flags = flags;
    P.rho = rho;
    P.mu = mu;
    P.dt = dt;
    for (int d = 0; d < 3; ++d) P.g[d] = d < dim ? g[d] : 0.0;
    auto fl = derivedFlags(P);
    auto* h = new orc_sys();
    if (dim == 2) {
        buildAbPSPG<2>(P, fl, pos, q_prev, q_iter, h->S);
        if (apply_bc) applyBCPSPG<2>(P, fl, pos, q_prev, dir_vals, h->S);
    } else {
        buildAbPSPG<3>(P, fl, pos, q_prev, q_iter, h->S);
        if (apply_bc) applyBCPSPG<3>(P, fl, pos, q_prev, dir_vals, h->S);
    }
    return h;
}
int64_t orc_sys_n(const orc_sys* h) { return h->S.n; }
int64_t orc_sys_nnz(const orc_sys* h) { return h->S.rowPtr.empty() ? 0 : h->S.rowPtr.back(); }
void orc_sys_copy(const orc_sys* h, int* row_ptr, int* col_idx, double* vals, double* b) {
    const auto& S = h->S;
    if (row_ptr) std::copy(S.rowPtr.begin(), S.rowPtr.end(), row_ptr);
    if (col_idx) std::copy(S.colIdx.begin(), S.colIdx.end(), col_idx);
    if (vals) std::copy(S.vals.begin(), S.vals.end(), vals);
    if (b) std::copy(S.b.begin(), S.b.end(), b);
}
void orc_sys_times(const orc_sys* h, double* out5) {
    out5[0] = h->S.tTriplets;
    out5[1] = h->S.tPush;
    out5[2] = h->S.tAssemble;
    out5[3] = h->S.tVector;
    out5[4] = h->S.tBC;
}
void orc_sys_free(orc_sys* h) { delete h; }

// Element-level access for closed-form pins: Ae, be of one element (row-major (d+1)^2 x (d+1)^2).
void orc_pspg_element(int dim, const double* pos_e, const double* qprev_e, const double* qiter_e, double rho, double mu,
                      const double* g, double dt, double* Ae_out, double* be_out, double* tau_out) {
    // pos_e: dim x (dim+1) field-blocked for a standalone element with nodes 0..dim
    Params P{};
    P.dim = dim;
    P.N = dim + 1;
    P.E = 1;
    P.rho = rho;
    P.mu = mu;
    P.dt = dt;
    for (int d = 0; d < 3; ++d) P.g[d] = d < dim ? g[d] : 0.0;
    int nodes[4] = {0, 1, 2, 3};
    if (dim == 2) {
        double Ae[9][9], be[9];
        MB<2> mb;
        pspgElement<2>(P, mb, pos_e, nodes, qprev_e, qiter_e, Ae, be);
        std::memcpy(Ae_out, Ae, sizeof(Ae));
        std::memcpy(be_out, be, sizeof(be));
        Geo geo;
        computeGeo<2>(pos_e, P.N, nodes, geo);
        *tau_out = tauPSPG<2>(P, geo, nodes, qiter_e);
    } else {
        double Ae[16][16], be[16];
        MB<3> mb;
        pspgElement<3>(P, mb, pos_e, nodes, qprev_e, qiter_e, Ae, be);
        std::memcpy(Ae_out, Ae, sizeof(Ae));
        std::memcpy(be_out, be, sizeof(be));
        Geo geo;
        computeGeo<3>(pos_e, P.N, nodes, geo);
        *tau_out = tauPSPG<3>(P, geo, nodes, qiter_e);
    }
}

// Element matrices (factor 1 / given) for invariant pins: M (rho), K (mu, incompressible ddev),
// D, C, L (factor 1), gradN, detJ.  Sizes: ne*ne, nd*nd, ne*nd, ne*nd, ne*ne, dim*ne.
void orc_element_matrices(int dim, const double* pos_e, double rho, double mu, double* M, double* K, double* Dm,
                          double* C, double* L, double* G, double* detJ, double* rin) {
    int nodes[4] = {0, 1, 2, 3};
    if (dim == 2) {
        MB<2> mb;
        Geo geo;
        computeGeo<2>(pos_e, 3, nodes, geo);
        double Gm[2][3], Bm[3][6], Mm[3][3], Km[6][6], Dd[3][6], Cc[3][6], Ll[3][3], dd[3][3];
        mb.gradN(geo, Gm);
        mb.B(Gm, Bm);
        mb.M(geo.detJ, [&](const double*) { return rho; }, Mm);
        ddevIncomp<2>(dd);
        mb.K(geo.detJ, mu, dd, Bm, Km);
        mb.Dm(geo.detJ, [](const double*) { return 1.0; }, Bm, Dd);
        mb.C(geo.detJ, Gm, Cc);
        mb.L(geo.detJ, 1.0, Gm, Ll);
        std::memcpy(M, Mm, sizeof(Mm));
        std::memcpy(K, Km, sizeof(Km));
        std::memcpy(Dm, Dd, sizeof(Dd));
        std::memcpy(C, Cc, sizeof(Cc));
        std::memcpy(L, Ll, sizeof(Ll));
        std::memcpy(G, Gm, sizeof(Gm));
        *detJ = geo.detJ;
        *rin = getRin<2>(pos_e, 3, nodes, geo.detJ);
    } else {
        MB<3> mb;
        Geo geo;
        computeGeo<3>(pos_e, 4, nodes, geo);
        double Gm[3][4], Bm[6][12], Mm[4][4], Km[12][12], Dd[4][12], Cc[4][12], Ll[4][4], dd[6][6];
        mb.gradN(geo, Gm);
        mb.B(Gm, Bm);
        mb.M(geo.detJ, [&](const double*) { return rho; }, Mm);
        ddevIncomp<3>(dd);
        mb.K(geo.detJ, mu, dd, Bm, Km);
        mb.Dm(geo.detJ, [](const double*) { return 1.0; }, Bm, Dd);
        mb.C(geo.detJ, Gm, Cc);
        mb.L(geo.detJ, 1.0, Gm, Ll);
        std::memcpy(M, Mm, sizeof(Mm));
        std::memcpy(K, Km, sizeof(Km));
        std::memcpy(Dm, Dd, sizeof(Dd));
        std::memcpy(C, Cc, sizeof(Cc));
        std::memcpy(L, Ll, sizeof(Ll));
        std::memcpy(G, Gm, sizeof(Gm));
        *detJ = geo.detJ;
        *rin = getRin<3>(pos_e, 4, nodes, geo.detJ);
    }
}

// Linear solve of an oracle system: kind 0 dense LU, 1 BiCGSTAB+Jacobi, 2 BiCGSTAB+ILU(0).
int orc_sys_solve(const orc_sys* h, int kind, double rtol, int maxit, double* x, int* iters, double* relres,
                  int nthreads) {
    setThreads(nthreads);
    SolveCfg cfg{kind, rtol, maxit};
    return linearSolve(h->S, cfg, x, iters, relres);
}

// CSR solve (for timing the Krylov stand-in on arbitrary systems).
int orc_csr_bicgstab(int64_t n, const int* rp, const int* ci, const double* v, const double* b, double* x, int prec,
                     double rtol, int maxit, int* iters, double* relres, int nthreads) {
    setThreads(nthreads);
    return bicgstab(n, rp, ci, v, b, x, prec, rtol, maxit, iters, relres);
}

// Whole PSPG time step (Picard).  stats_out (16 doubles): picardIters, krylovIters,
// assemblies, finalRes, maxRelres, tAssemble, tSolve, tUpdate, tResidual, tTriplets, tPush,
// tSetFromTriplets, tVector, tBC.
int orc_pspg_step(int dim, int64_t N, int64_t E, const int* conn, const uint8_t* flags, double* pos, double* q,
                  const double* dir_vals, double rho, double mu, const double* g, double dt, int res_kind,
                  double min_res, int max_iter, int solver_kind, double rtol, int krylov_maxit, double* stats_out,
                  int nthreads, const double* dir_affine) {
    setThreads(nthreads);
    Params P{};
    P.dim = dim;
    P.N = N;
    P.E = E;
    P.conn = conn;
    P.flags = flags;
    P.rho = rho;
    P.mu = mu;
    P.dt = dt;
    for (int d = 0; d < 3; ++d) P.g[d] = d < dim ? g[d] : 0.0;
    SolveCfg cfg{solver_kind, rtol, krylov_maxit};
    StepStats st;
    int s = (dim == 2) ? pspgStep<2>(P, pos, q, dir_vals, res_kind, min_res, max_iter, cfg, st, dir_affine)
                       : pspgStep<3>(P, pos, q, dir_vals, res_kind, min_res, max_iter, cfg, st, dir_affine);
    if (stats_out) {
        double o[14] = {(double)st.picardIters, (double)st.krylovIters, (double)st.assemblies, st.finalRes,
                        st.maxRelres,           st.tAssemble,           st.tSolve,          st.tUpdate,
                        st.tResidual,           st.tTriplets,           st.tPush,           st.tSetFromTriplets,
                        st.tVector,             st.tBC};
        std::memcpy(stats_out, o, sizeof(o));
    }
    return s;
}

// One WC CDS_dpdt step; state arrays are updated in place.  times_out: [tCont, tMom].
int orc_wc_step(int dim, int64_t N, int64_t E, const int* conn, const uint8_t* flags, double* pos, double* v,
                double* p, double* rho, double* a, const double* dir_vals, double mu, const double* g, double K0,
                double K0p, double rho_star, int meduri, double dt, double* times_out, int nthreads) {
    setThreads(nthreads);
    WCParams W{};
    W.dim = dim;
    W.N = N;
    W.E = E;
    W.conn = conn;
    W.flags = flags;
    W.mu = mu;
    for (int d = 0; d < 3; ++d) W.g[d] = d < dim ? g[d] : 0.0;
    W.K0 = K0;
    W.K0p = K0p;
    W.rhoStar = rho_star;
    W.meduri = meduri;
    double tc = 0, tm = 0;
    if (dim == 2)
        wcStep<2>(W, dt, pos, v, p, rho, a, dir_vals, &tc, &tm);
    else
        wcStep<3>(W, dt, pos, v, p, rho, a, dir_vals, &tc, &tm);
    if (times_out) {
        times_out[0] = tc;
        times_out[1] = tm;
    }
    return 0;
}

double orc_wc_next_dt(int dim, int64_t N, int64_t E, const int* conn, const double* pos, const double* v,
                      const double* p, const double* rho, double mu, double K0, double K0p, double security_coeff,
                      double max_dt, int nthreads) {
    setThreads(nthreads);
    WCParams W{};
    W.dim = dim;
    W.N = N;
    W.E = E;
    W.conn = conn;
    W.mu = mu;
    W.K0 = K0;
    W.K0p = K0p;
    return dim == 2 ? wcNextDT<2>(W, pos, v, p, rho, security_coeff, max_dt)
                    : wcNextDT<3>(W, pos, v, p, rho, security_coeff, max_dt);
}

// Global "mass" of a problem: ProbIncompNewton::getGlobalWrittableData
// (IncompNewton/Problem.cpp:173-190) sums the element sizes Element::getSize = detJ * refSize
// (Element.cpp:159-162); ProbWCompNewton::getGlobalWrittableData (WCompNewton/Problem.cpp:216-240)
// weights each size by the mean density of the element's nodes (rho non-null).  Elements in
// index order, as the reference loops.
double orc_mass(int dim, int64_t N, int64_t E, const int* conn, const double* pos, const double* rho) {
    const int ne = dim + 1;
    double mass = 0;
    for (int64_t e = 0; e < E; ++e) {
        const int* nodes = conn + e * ne;
        Geo g;
        double size;
        if (dim == 2) {
            computeGeo<2>(pos, N, nodes, g);
            size = g.detJ * Quad<2>::refSize;
        } else {
            computeGeo<3>(pos, N, nodes, g);
            size = g.detJ * Quad<3>::refSize;
        }
        if (rho) {
            double middle = 0;
            for (int n = 0; n < ne; ++n) middle += rho[nodes[n]];
            middle /= static_cast<double>(ne);
            mass += middle * size;
        } else {
            mass += size;
        }
    }
    return mass;
}

// MinMaxExtractor::update (MinMaxExtractor.cpp:35-55): the smallest (want_max = 0) or largest
// coordinate `coordinate` over the nodes that are neither bound nor free; the largest starts
// from 0, as the reference's does, the smallest from DBL_MAX.
double orc_minmax(int dim, int64_t N, const uint8_t* flags, const double* pos, int coordinate, int want_max) {
    (void)dim;
    double v = want_max ? 0.0 : std::numeric_limits<double>::max();
    for (int64_t n = 0; n < N; ++n) {
        if ((flags[n] & BOUND) || (flags[n] & FREE)) continue;
        const double x = pos[n + coordinate * N];
        v = want_max ? std::max(v, x) : std::min(v, x);
    }
    return v;
}

// Residual callback on given iterates (kind 0 U, 1 U_P, 2 Ax_f).  flags are used as given
// (FREE included); for Ax_f the CSR system (rp, ci, vals, b) of n = (dim+1) N rows is used.
double orc_picard_residual(int dim, int64_t N, const uint8_t* flags, int kind, const int* rp, const int* ci,
                           const double* vals, const double* b, const double* q_iter, const double* q_iter_prev) {
    CsrSys S;
    const int64_t n = (dim + 1) * N;
    if (kind == 2) {
        S.n = n;
        S.rowPtr.assign(rp, rp + n + 1);
        S.colIdx.assign(ci, ci + rp[n]);
        S.vals.assign(vals, vals + rp[n]);
        S.b.assign(b, b + n);
    }
    return picardResidual(dim, N, flags, kind, S, q_iter, q_iter_prev);
}

// The two scattered WC systems of one step at the given state (no position update, no BC):
// ML, F0 (N each) of the continuity equation and Mv, F (dim x N each) of the momentum
// equation; rho is the density the momentum system reads.
void orc_wc_parts(int dim, int64_t N, int64_t E, const int* conn, const double* pos, const double* v,
                  const double* p, const double* rho, double mu, const double* g, double K0, double K0p, int meduri,
                  double dt, double* ML, double* F0, double* Mv, double* F, int nthreads) {
    setThreads(nthreads);
    WCParams W{};
    W.dim = dim;
    W.N = N;
    W.E = E;
    W.conn = conn;
    W.mu = mu;
    for (int d = 0; d < 3; ++d) W.g[d] = d < dim ? g[d] : 0.0;
    W.K0 = K0;
    W.K0p = K0p;
    W.meduri = meduri;
    if (dim == 2) {
        wcContAssemble<2>(W, dt, pos, v, p, ML, F0);
        wcMomAssemble<2>(W, pos, v, p, rho, Mv, F);
    } else {
        wcContAssemble<3>(W, dt, pos, v, p, ML, F0);
        wcMomAssemble<3>(W, pos, v, p, rho, Mv, F);
    }
}

// Material law of the PSPG entry points that follow (problem 0 Newtonian, 1 Bingham).
void orc_set_material(int problem, double tau0, double m_reg) {
    g_mat.problem = problem;
    g_mat.tau0 = tau0;
    g_mat.mReg = m_reg;
}

// Surface tension: F boundary facets (nodes F x dim, element F), gamma and DgammaDT (F = 0 clears)
void orc_set_facets(int64_t F, int dim, const int* nodes, const int* elem, double gamma, double dgdT) {
    g_fst.F = F;
    g_fst.nodes.assign(nodes, nodes + F * dim);
    g_fst.elem.assign(elem, elem + F);
    g_fst.gamma = gamma;
    g_fst.dgdT = dgdT;
}

// Boussinesq factors: thermal expansion alpha, reference temperature Tr, node temperatures
// (N values, copied; NULL clears)
void orc_set_boussinesq(double alpha, double Tr, const double* T, int64_t N) {
    g_mat.alpha = alpha;
    g_mat.Tr = Tr;
    if (T) {
        g_T.assign(T, T + N);
        g_mat.T = g_T.data();
    } else {
        g_T.clear();
        g_mat.T = nullptr;
    }
}

}  // extern "C"

extern "C" {
// The heat system (m_buildAb + m_applyBC) of one step from T (the previous step's node states):
// CSR of N rows and the rhs.
orc_sys* orc_heat_assemble(int dim, int64_t N, int64_t E, const int* conn, const uint8_t* flags, const double* pos,
                           const double* T, const double* Tbc, double rho, double cv, double k, double h,
                           double Tinf, double epsRad, double dt, int64_t F, const int* facets, const int* fout,
                           const uint8_t* fkind, const double* qgp) {
    HeatParams P;
    P.dim = dim;
    P.N = N;
    P.E = E;
    P.F = F;
    P.conn = conn;
    P.flags = flags;
    P.rho = rho;
    P.cv = cv;
    P.k = k;
    P.h = h;
    P.Tinf = Tinf;
    P.epsRad = epsRad;
    P.dt = dt;
    P.facets = facets;
    P.fout = fout;
    P.fkind = fkind;
    P.qgp = qgp;
    auto* o = new orc_sys;
    if (dim == 2)
        heatSystem<2>(P, pos, T, Tbc, o->S);
    else
        heatSystem<3>(P, pos, T, Tbc, o->S);
    return o;
}

// One HeatEqIncompNewton::solve without phase change: the system, the CG solve from T, the new
// node states in T.  Returns 0, 1 (not converged) or 2 (NaN).
int orc_heat_step(int dim, int64_t N, int64_t E, const int* conn, const uint8_t* flags, const double* pos, double* T,
                  const double* Tbc, double rho, double cv, double k, double h, double Tinf, double epsRad, double dt,
                  int64_t F, const int* facets, const int* fout, const uint8_t* fkind, const double* qgp, double rtol,
                  int maxit, int* iters, double* relres) {
    orc_sys* o = orc_heat_assemble(dim, N, E, conn, flags, pos, T, Tbc, rho, cv, k, h, Tinf, epsRad, dt, F, facets,
                                   fout, fkind, qgp);
    std::vector<double> x(T, T + N);
    const int st = pcg(N, o->S.rowPtr.data(), o->S.colIdx.data(), o->S.vals.data(), o->S.b.data(), x.data(), rtol,
                       maxit, iters, relres);
    delete o;
    if (st == 0) std::copy(x.begin(), x.end(), T);
    return st;
}

// getQN / getSGamma of every facet from the node temperatures T (F x dim each).
void orc_heat_facet_terms(int dim, int64_t N, const double* pos, const double* T, int64_t F, const int* facets,
                          const int* fout, const double* qgp, double h, double Tinf, double epsRad, double* qn,
                          double* sgh, double* sgr) {
    HeatParams P;
    P.dim = dim;
    P.N = N;
    P.F = F;
    P.facets = facets;
    P.fout = fout;
    P.qgp = qgp;
    P.h = h;
    P.Tinf = Tinf;
    P.epsRad = epsRad;
    for (int64_t f = 0; f < F; ++f) {
        if (dim == 2)
            heatFacetTerms<2>(P, pos, T, f, qn + f * 2, sgh + f * 2, sgr + f * 2);
        else
            heatFacetTerms<3>(P, pos, T, f, qn + f * 3, sgh + f * 3, sgr + f * 3);
    }
}
}  // extern "C"

extern "C" {
// One FracStep time step; q = (v, p) field-blocked, (dim+1) N; resKind 0 "U", 1 "U_P".
// stats5: {Picard iterations, CG iterations of the three solves, final residual}.
int orc_frac_step(int dim, int64_t N, int64_t E, const int* conn, const uint8_t* flags, double* pos, double* q,
                  const double* dirVals, double rho, double mu, const double* g, double dt, double gammaFS,
                  int resKind, double minRes, int maxIter, double rtol, int cgMaxit, double* stats5) {
    Params P{};
    P.dim = dim;
    P.N = N;
    P.E = E;
    P.conn = conn;
    P.flags = flags;
    P.rho = rho;
    P.mu = mu;
    for (int k = 0; k < 3; ++k) P.g[k] = g[k];
    P.dt = dt;
    FracStats st;
    const int r = dim == 2 ? fracStep<2>(P, pos, q, dirVals, gammaFS, resKind, minRes, maxIter, rtol, cgMaxit, st)
                           : fracStep<3>(P, pos, q, dirVals, gammaFS, resKind, minRes, maxIter, rtol, cgMaxit, st);
    if (stats5) {
        stats5[0] = st.picardIters;
        for (int k = 0; k < 3; ++k) stats5[1 + k] = st.cgIters[k];
        stats5[4] = st.finalRes;
    }
    return r;
}
}  // extern "C"
