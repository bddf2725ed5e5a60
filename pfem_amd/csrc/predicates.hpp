// Exact geometric predicates for the host mesher (orientation, in-circle, in-sphere) with the
// lexicographic symbolic perturbation that makes the Delaunay triangulation of any point set
// unique.
//
// The reference triangulates with CGAL's Epick kernel (exact predicates, srcs/mesh/Mesh2D.cpp:19,
// Mesh3D.cpp:20) whose Delaunay insertion resolves co-circular / co-spherical points by symbolic
// perturbation (Devillers & Teillaud, "Perturbations for Delaunay and weighted Delaunay 3D
// triangulations", Comput. Geom. 44(3), 2011).  The perturbation used here is that paper's: the
// lifted coordinate |p|^2 of every point is raised by an infinitesimal whose order follows the
// lexicographic (x, y[, z]) rank of the point, the largest point getting the dominant one.  When the
// exact determinant is zero, the sign of the perturbed determinant is the sign of the coefficient
// of the most dominant point whose coefficient is nonzero; the coefficient of a cell vertex v is the
// orientation of the cell with v replaced by the query point, that of the query point is minus the
// cell's orientation (always "outside").  The triangulation this defines is the regular
// triangulation of the perturbed weights, unique, so any correct insertion order yields it.
//
// Arithmetic: a floating-point evaluation with a forward error bound (the determinant's
// permanent times a constant derived for each formula below) decides the sign when it can; the
// rest are evaluated exactly with floating-point expansions (Shewchuk, "Adaptive precision
// floating-point arithmetic and fast robust geometric predicates", DCG 18, 1997: error-free
// two-sum and two-product, expansion sum by merge and two-sum sweep, scaling by a double).
// The bounds here are looser than that paper's tight ones (about twice their leading constants):
// a looser bound only sends more cases to the exact path.
//
// Sign conventions (all return -1, 0, +1):
//   orient2(a,b,c)       sign of det[b-a; c-a]            (+: a, b, c counterclockwise)
//   incircle2(a,b,c,d)   for counterclockwise a,b,c     (+: d strictly inside their circle)
//   orient3(a,b,c,d)     sign of det[b-a; c-a; d-a]       (+: the element detJ > 0 orientation)
//   insphere3(a,b,c,d,e) for orient3(a,b,c,d) > 0       (+: e strictly inside their sphere)
// The *_sos variants never return 0 for distinct points.
#pragma once

#include <algorithm>
#include <cmath>
#include <vector>

namespace pfem {
namespace geom {

constexpr double kEps = 1.1102230246251565e-16;  // 2^-53, unit roundoff of round-to-nearest

// ---- error-free transformations -------------------------------------------------------------
inline void two_sum(double a, double b, double& s, double& e)
{
    s = a + b;
    const double bb = s - a;
    const double ab = s - bb;
    e = (a - ab) + (b - bb);
}

inline void two_prod(double a, double b, double& p, double& e)
{
    p = a * b;
    e = std::fma(a, b, -p);
}

// ---- expansions: nonoverlapping components, increasing magnitude, no zeros ---------------------
using Ex = std::vector<double>;

inline int ex_sign(const Ex& e) { return e.empty() ? 0 : (e.back() > 0 ? 1 : -1); }

inline Ex ex_of(double a)
{
    Ex r;
    if (a != 0) r.push_back(a);
    return r;
}

inline Ex ex_diff(double a, double b)
{
    double s, e;
    two_sum(a, -b, s, e);
    Ex r;
    if (e != 0) r.push_back(e);
    if (s != 0) r.push_back(s);
    return r;
}

// e + f: merge by magnitude, then one two-sum sweep carrying the running sum upward
inline Ex ex_add(const Ex& e, const Ex& f)
{
    if (e.empty()) return f;
    if (f.empty()) return e;
    Ex g(e.size() + f.size());
    std::merge(e.begin(), e.end(), f.begin(), f.end(), g.begin(),
               [](double x, double y) { return std::fabs(x) < std::fabs(y); });
    Ex h;
    h.reserve(g.size());
    double q = g[0];
    for (size_t i = 1; i < g.size(); ++i) {
        double s, t;
        two_sum(q, g[i], s, t);
        if (t != 0) h.push_back(t);
        q = s;
    }
    if (q != 0 || h.empty()) h.push_back(q);
    if (h.size() == 1 && h[0] == 0) h.clear();
    return h;
}

inline Ex ex_neg(Ex e)
{
    for (double& x : e) x = -x;
    return e;
}

inline Ex ex_sub(const Ex& e, const Ex& f) { return ex_add(e, ex_neg(f)); }

// e * b
inline Ex ex_scale(const Ex& e, double b)
{
    Ex h;
    if (e.empty() || b == 0) return h;
    h.reserve(2 * e.size());
    double q, t;
    two_prod(e[0], b, q, t);
    if (t != 0) h.push_back(t);
    for (size_t i = 1; i < e.size(); ++i) {
        double p1, p0, s, u;
        two_prod(e[i], b, p1, p0);
        two_sum(q, p0, s, u);
        if (u != 0) h.push_back(u);
        two_sum(p1, s, q, u);
        if (u != 0) h.push_back(u);
    }
    if (q != 0) h.push_back(q);
    return h;
}

inline Ex ex_mul(const Ex& e, const Ex& f)
{
    Ex acc;
    for (double b : f) acc = ex_add(acc, ex_scale(e, b));
    return acc;
}

// ---- determinants over expansions (the exact path) -------------------------------------------
inline Ex ex_det2(const Ex& a, const Ex& b, const Ex& c, const Ex& d)  // | a b ; c d |
{
    return ex_sub(ex_mul(a, d), ex_mul(b, c));
}

inline Ex ex_det3(const Ex* m)  // 3x3 row-major
{
    Ex t0 = ex_mul(m[0], ex_det2(m[4], m[5], m[7], m[8]));
    Ex t1 = ex_mul(m[1], ex_det2(m[3], m[5], m[6], m[8]));
    Ex t2 = ex_mul(m[2], ex_det2(m[3], m[4], m[6], m[7]));
    return ex_add(ex_sub(t0, t1), t2);
}

// ---- orient2 ---------------------------------------------------------------------------------
inline int orient2_exact(const double* a, const double* b, const double* c)
{
    Ex bx = ex_diff(b[0], a[0]), by = ex_diff(b[1], a[1]);
    Ex cx = ex_diff(c[0], a[0]), cy = ex_diff(c[1], a[1]);
    return ex_sign(ex_det2(bx, by, cx, cy));
}

inline int orient2(const double* a, const double* b, const double* c)
{
    const double bx = b[0] - a[0], by = b[1] - a[1], cx = c[0] - a[0], cy = c[1] - a[1];
    const double l = bx * cy, r = by * cx;
    const double det = l - r;
    // differences and products each carry one rounding: |error| <= (3 eps + O(eps^2)) (|l| + |r|)
    // plus the final subtraction's eps |det|
    const double bound = 6.0 * kEps * (std::fabs(l) + std::fabs(r));
    if (det > bound) return 1;
    if (-det > bound) return -1;
    return orient2_exact(a, b, c);
}

// ---- incircle2 -------------------------------------------------------------------------------
inline int incircle2_exact(const double* a, const double* b, const double* c, const double* d)
{
    Ex m[9];
    const double* p[3] = {a, b, c};
    for (int i = 0; i < 3; ++i) {
        Ex x = ex_diff(p[i][0], d[0]), y = ex_diff(p[i][1], d[1]);
        m[3 * i + 2] = ex_add(ex_mul(x, x), ex_mul(y, y));
        m[3 * i + 0] = std::move(x);
        m[3 * i + 1] = std::move(y);
    }
    return ex_sign(ex_det3(m));
}

inline int incircle2(const double* a, const double* b, const double* c, const double* d)
{
    const double adx = a[0] - d[0], ady = a[1] - d[1];
    const double bdx = b[0] - d[0], bdy = b[1] - d[1];
    const double cdx = c[0] - d[0], cdy = c[1] - d[1];
    const double al = adx * adx + ady * ady, bl = bdx * bdx + bdy * bdy, cl = cdx * cdx + cdy * cdy;
    const double m_a = bdx * cdy - bdy * cdx, m_b = cdx * ady - cdy * adx, m_c = adx * bdy - ady * bdx;
    const double det = al * m_a + bl * m_b + cl * m_c;
    const double perm = al * (std::fabs(bdx * cdy) + std::fabs(bdy * cdx)) +
                        bl * (std::fabs(cdx * ady) + std::fabs(cdy * adx)) +
                        cl * (std::fabs(adx * bdy) + std::fabs(ady * bdx));
    // lift: 4 eps relative; 2x2 minors: 4 eps of their permanent; product and the two sums:
    // about 11 eps of the permanent in total
    const double bound = 24.0 * kEps * perm;
    if (det > bound) return 1;
    if (-det > bound) return -1;
    return incircle2_exact(a, b, c, d);
}

// ---- orient3 ---------------------------------------------------------------------------------
inline int orient3_exact(const double* a, const double* b, const double* c, const double* d)
{
    Ex m[9];
    const double* p[3] = {b, c, d};
    for (int i = 0; i < 3; ++i)
        for (int k = 0; k < 3; ++k) m[3 * i + k] = ex_diff(p[i][k], a[k]);
    return ex_sign(ex_det3(m));
}

inline int orient3(const double* a, const double* b, const double* c, const double* d)
{
    const double bx = b[0] - a[0], by = b[1] - a[1], bz = b[2] - a[2];
    const double cx = c[0] - a[0], cy = c[1] - a[1], cz = c[2] - a[2];
    const double dx = d[0] - a[0], dy = d[1] - a[1], dz = d[2] - a[2];
    const double m0 = cy * dz - cz * dy, m1 = cx * dz - cz * dx, m2 = cx * dy - cy * dx;
    const double det = bx * m0 - by * m1 + bz * m2;
    const double perm = std::fabs(bx) * (std::fabs(cy * dz) + std::fabs(cz * dy)) +
                        std::fabs(by) * (std::fabs(cx * dz) + std::fabs(cz * dx)) +
                        std::fabs(bz) * (std::fabs(cx * dy) + std::fabs(cy * dx));
    // differences 1 eps, minors 3-4 eps, products and two sums: about 7 eps of the permanent
    const double bound = 16.0 * kEps * perm;
    if (det > bound) return 1;
    if (-det > bound) return -1;
    return orient3_exact(a, b, c, d);
}

// ---- insphere3 -------------------------------------------------------------------------------
// value = -det[[a-e, |a-e|^2]; [b-e, .]; [c-e, .]; [d-e, .]] (4x4), positive for e inside when
// orient3(a,b,c,d) > 0
inline int insphere3_exact(const double* a, const double* b, const double* c, const double* d, const double* e)
{
    const double* p[4] = {a, b, c, d};
    Ex x[4][3], l[4];
    for (int i = 0; i < 4; ++i) {
        for (int k = 0; k < 3; ++k) x[i][k] = ex_diff(p[i][k], e[k]);
        l[i] = ex_add(ex_add(ex_mul(x[i][0], x[i][0]), ex_mul(x[i][1], x[i][1])), ex_mul(x[i][2], x[i][2]));
    }
    // expand along the lift column: det = sum_i (-1)^(i+3) l_i * M_i, M_i the 3x3 minor of rows != i
    Ex det;
    for (int i = 0; i < 4; ++i) {
        Ex m[9];
        int r = 0;
        for (int j = 0; j < 4; ++j) {
            if (j == i) continue;
            for (int k = 0; k < 3; ++k) m[3 * r + k] = x[j][k];
            ++r;
        }
        Ex t = ex_mul(l[i], ex_det3(m));
        det = ((i + 3) % 2 == 0) ? ex_add(det, t) : ex_sub(det, t);
    }
    return -ex_sign(det);
}

inline int insphere3(const double* a, const double* b, const double* c, const double* d, const double* e)
{
    const double* p[4] = {a, b, c, d};
    double x[4][3], l[4];
    for (int i = 0; i < 4; ++i) {
        for (int k = 0; k < 3; ++k) x[i][k] = p[i][k] - e[k];
        l[i] = x[i][0] * x[i][0] + x[i][1] * x[i][1] + x[i][2] * x[i][2];
    }
    // 2x2 minors of the (x, y) columns for every row pair, then 3x3 minors with z
    auto m2 = [&](int i, int j) { return x[i][0] * x[j][1] - x[j][0] * x[i][1]; };
    auto p2 = [&](int i, int j) { return std::fabs(x[i][0] * x[j][1]) + std::fabs(x[j][0] * x[i][1]); };
    // minor of rows (i, j, k) (ascending) = z_i m2(j,k) - z_j m2(i,k) + z_k m2(i,j)
    auto m3 = [&](int i, int j, int k) { return x[i][2] * m2(j, k) - x[j][2] * m2(i, k) + x[k][2] * m2(i, j); };
    auto p3 = [&](int i, int j, int k) {
        return std::fabs(x[i][2]) * p2(j, k) + std::fabs(x[j][2]) * p2(i, k) + std::fabs(x[k][2]) * p2(i, j);
    };
    const double M0 = m3(1, 2, 3), M1 = m3(0, 2, 3), M2 = m3(0, 1, 3), M3 = m3(0, 1, 2);
    // det = -l0 M0 + l1 M1 - l2 M2 + l3 M3  (cofactor signs (-1)^(i+4) with rows 1-based)
    const double det = -l[0] * M0 + l[1] * M1 - l[2] * M2 + l[3] * M3;
    const double perm = l[0] * p3(1, 2, 3) + l[1] * p3(0, 2, 3) + l[2] * p3(0, 1, 3) + l[3] * p3(0, 1, 2);
    // lifts 5 eps relative, 3x3 minors about 7 eps of their permanents, products and three sums:
    // about 16 eps of the permanent
    const double bound = 32.0 * kEps * perm;
    if (det > bound) return -1;
    if (-det > bound) return 1;
    return insphere3_exact(a, b, c, d, e);
}

// ---- symbolic perturbation -------------------------------------------------------------------
template <int D>
inline bool lex_less(const double* p, const double* q)
{
    for (int k = 0; k < D; ++k) {
        if (p[k] < q[k]) return true;
        if (p[k] > q[k]) return false;
    }
    return false;
}

// a, b, c counterclockwise; never 0 for four distinct points
inline int incircle2_sos(const double* a, const double* b, const double* c, const double* d)
{
    const int s = incircle2(a, b, c, d);
    if (s != 0) return s;
    const double* pts[4] = {a, b, c, d};
    int order[4] = {0, 1, 2, 3};
    std::sort(order, order + 4, [&](int i, int j) { return lex_less<2>(pts[j], pts[i]); });  // descending
    for (int r = 0; r < 4; ++r) {
        const int v = order[r];
        if (v == 3) return -1;  // the query point's own coefficient: -orient2(a,b,c) < 0
        const double* q[3] = {a, b, c};
        q[v] = d;
        const int o = orient2(q[0], q[1], q[2]);
        if (o != 0) return o;
    }
    return -1;
}

// orient3(a,b,c,d) > 0; never 0 for five distinct points
inline int insphere3_sos(const double* a, const double* b, const double* c, const double* d, const double* e)
{
    const int s = insphere3(a, b, c, d, e);
    if (s != 0) return s;
    const double* pts[5] = {a, b, c, d, e};
    int order[5] = {0, 1, 2, 3, 4};
    std::sort(order, order + 5, [&](int i, int j) { return lex_less<3>(pts[j], pts[i]); });
    for (int r = 0; r < 5; ++r) {
        const int v = order[r];
        if (v == 4) return -1;
        const double* q[4] = {a, b, c, d};
        q[v] = e;
        const int o = orient3(q[0], q[1], q[2], q[3]);
        if (o != 0) return o;
    }
    return -1;
}

}  // namespace geom
}  // namespace pfem
