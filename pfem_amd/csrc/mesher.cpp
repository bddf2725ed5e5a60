// Host remesher: Delaunay triangulation, alpha-shape filter and node addition / removal of the
// reference's Lagrangian mesh (srcs/mesh/Mesh.cpp:27-331,580-760,919-1017, Mesh2D.cpp:32-251,
// Mesh3D.cpp:36-286), without CGAL.
//
// * Triangulation: Bowyer-Watson insertion over a triangulation closed by one infinite vertex,
//   points in biased-randomized rounds each sorted along a Hilbert curve, located by a visibility
//   walk from the last created simplex.  Predicates are exact with the lexicographic symbolic
//   perturbation (predicates.hpp), so the result is THE perturbed Delaunay triangulation that
//   CGAL's Delaunay_triangulation_2/_3 also builds: the element set matches the reference's for
//   any input, co-circular lattices included.  The reference's element ORDER is CGAL's internal
//   iteration order; here elements are sorted by their canonical node tuple (smallest node first,
//   orientation kept), which fixes the numbering of the nodes addNodes appends.
// * Alpha shape: a simplex is interior when its squared circumradius is <= (alpha hchar)^2 (the
//   alpha-shape parameter the reference passes, Mesh2D.cpp:56, Mesh3D.cpp:61); interior simplices
//   are kept unless Mesh2D.cpp:58-118 / Mesh3D.cpp:63-128 delete them; free-surface facets are the
//   faces between an interior simplex (kept) and a non-interior one (Mesh2D.cpp:192-242,
//   Mesh3D.cpp:215-282).
// * The mesher is host code: it runs on the CPU between GPU steps (the reference remeshes after
//   every successful PSPG step, IncompNewton/Solver.cpp:236-246), one mesher per case, so several
//   cases' remeshes run on separate host threads while the GPU steps others.
#include <algorithm>
#include <array>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <numeric>
#include <random>
#include <stdexcept>
#include <type_traits>
#include <string>
#include <vector>

#include "predicates.hpp"
#include "mesher.hpp"

namespace pfem {
namespace mesh {

using geom::incircle2_sos;
using geom::insphere3_sos;
using geom::orient2;
using geom::orient3;

// ---------------------------------------------------------------------------------------------
// Hilbert keys (Skilling, "Programming the Hilbert curve", AIP Conf. Proc. 707, 2004: the axes
// are transformed in place to the curve's transposed index, whose bits interleave to the key)
static uint64_t hilbert_key(uint32_t* X, int n, int bits)
{
    const uint32_t M = 1u << (bits - 1);
    for (uint32_t Q = M; Q > 1; Q >>= 1) {
        const uint32_t P = Q - 1;
        for (int i = 0; i < n; ++i) {
            if (X[i] & Q) {
                X[0] ^= P;
            } else {
                const uint32_t t = (X[0] ^ X[i]) & P;
                X[0] ^= t;
                X[i] ^= t;
            }
        }
    }
    for (int i = 1; i < n; ++i) X[i] ^= X[i - 1];
    uint32_t t = 0;
    for (uint32_t Q = M; Q > 1; Q >>= 1)
        if (X[n - 1] & Q) t ^= Q - 1;
    for (int i = 0; i < n; ++i) X[i] ^= t;
    uint64_t key = 0;
    for (int b = bits - 1; b >= 0; --b)
        for (int i = 0; i < n; ++i) key = (key << 1) | ((X[i] >> b) & 1u);
    return key;
}

// insertion order: a seeded shuffle split into rounds of doubling size (the last holds half the
// points), each round sorted along the Hilbert curve of the whole bounding box
template <int D>
static std::vector<int> brio_order(const double* pts, const std::vector<int>& ids)
{
    const int n = (int)ids.size();
    std::vector<int> order(ids);
    std::mt19937_64 rng(0x5eed2024ULL);
    std::shuffle(order.begin(), order.end(), rng);
    double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
    for (int i : ids)
        for (int k = 0; k < D; ++k) {
            lo[k] = std::min(lo[k], pts[D * i + k]);
            hi[k] = std::max(hi[k], pts[D * i + k]);
        }
    const int bits = D == 2 ? 31 : 21;
    const double cells = double((1u << bits) - 1);
    double span = 0;
    for (int k = 0; k < D; ++k) span = std::max(span, hi[k] - lo[k]);
    const double scale = span > 0 ? cells / span : 0;
    std::vector<int> bounds;
    for (int e = n; e > 64; e /= 2) bounds.push_back(e);
    bounds.push_back(0);
    std::reverse(bounds.begin(), bounds.end());
    if (bounds.back() != n) bounds.push_back(n);
    std::vector<std::pair<uint64_t, int>> tmp;
    for (size_t r = 0; r + 1 < bounds.size(); ++r) {
        tmp.clear();
        for (int j = bounds[r]; j < bounds[r + 1]; ++j) {
            const int i = order[j];
            uint32_t X[3] = {0, 0, 0};
            for (int k = 0; k < D; ++k) X[k] = (uint32_t)std::min(cells, (pts[D * i + k] - lo[k]) * scale);
            tmp.push_back({hilbert_key(X, D, bits), i});
        }
        std::sort(tmp.begin(), tmp.end());
        for (int j = bounds[r]; j < bounds[r + 1]; ++j) order[j] = tmp[j - bounds[r]].second;
    }
    return order;
}

// ---------------------------------------------------------------------------------------------
// Triangulation<D>: simplices of D+1 vertices with the neighbour across the face opposite each.
// Finite simplices are positively oriented (orient2 / orient3 > 0).  The face opposite slot i is
// listed by FACE[i] in the orientation that sees vertex i on its positive side; a simplex holding
// the infinite vertex at slot i is the region beyond its finite face.
template <int D>
struct Face;
template <>
struct Face<2> {
    static constexpr int T[3][2] = {{1, 2}, {2, 0}, {0, 1}};
};
template <>
struct Face<3> {
    static constexpr int T[4][3] = {{1, 3, 2}, {0, 2, 3}, {0, 3, 1}, {0, 1, 2}};
};
constexpr int Face<2>::T[3][2];
constexpr int Face<3>::T[4][3];

template <int D>
class Triangulation {
public:
    static constexpr int NV = D + 1;
    struct Simplex {
        int v[NV];
        int n[NV];
    };

    // pts: n x D (row per point).  Returns the ids that were not inserted (exact duplicates).
    std::vector<int> build(const double* pts, int npts)
    {
        P_ = pts;
        inf_ = npts;
        S_.clear();
        free_.clear();
        std::vector<int> ids(npts);
        std::iota(ids.begin(), ids.end(), 0);
        std::vector<int> order = brio_order<D>(pts, ids);
        std::vector<int> skipped;
        if (!init_simplex(order)) throw std::runtime_error("the nodes do not span a " + std::to_string(D) + "D volume");
        rng_.seed(12345);
        for (int i : order) {
            if (i == init_[0] || i == init_[1] || i == init_[2] || (D == 3 && i == init_[3])) continue;
            if (!insert(i)) skipped.push_back(i);
        }
        return skipped;
    }

    const std::vector<Simplex>& simplices() const { return S_; }
    bool alive(int s) const { return S_[s].v[0] >= 0; }
    bool infinite(int s) const
    {
        for (int k = 0; k < NV; ++k)
            if (S_[s].v[k] == inf_) return true;
        return false;
    }
    int inf() const { return inf_; }

private:
    const double* P_ = nullptr;
    int inf_ = 0;
    std::vector<Simplex> S_;
    std::vector<int> free_;
    std::vector<uint32_t> mark_;
    uint32_t stamp_ = 0;
    int last_ = 0;
    int init_[4] = {-1, -1, -1, -1};
    std::mt19937 rng_;
    std::vector<int> cavity_, bfaces_, created_;
    // ridge hash (linking new simplices): open addressing; an entry is live when its generation
    // is the current one, so a new insertion starts from an empty table without clearing it
    struct HEntry {
        uint64_t key;
        int val;
        uint32_t gen;
    };
    std::vector<HEntry> ht_;
    uint32_t gen_ = 0;

    const double* pt(int v) const { return P_ + (size_t)D * v; }

    int orient(const int* f, const double* p) const
    {
        if constexpr (D == 2)
            return orient2(pt(f[0]), pt(f[1]), p);
        else
            return orient3(pt(f[0]), pt(f[1]), pt(f[2]), p);
    }

    // face opposite slot i of simplex s, oriented to see v[i] on its positive side
    void face(int s, int i, int* f) const
    {
        for (int k = 0; k < D; ++k) f[k] = S_[s].v[Face<D>::T[i][k]];
    }

    int inf_slot(int s) const
    {
        for (int k = 0; k < NV; ++k)
            if (S_[s].v[k] == inf_) return k;
        return -1;
    }

    bool conflict_finite(int s, const double* p) const
    {
        const int* v = S_[s].v;
        if constexpr (D == 2)
            return incircle2_sos(pt(v[0]), pt(v[1]), pt(v[2]), p) > 0;
        else
            return insphere3_sos(pt(v[0]), pt(v[1]), pt(v[2]), pt(v[3]), p) > 0;
    }

    bool conflict(int s, const double* p) const
    {
        const int i = inf_slot(s);
        if (i < 0) return conflict_finite(s, p);
        int f[D];
        face(s, i, f);
        const int o = orient(f, p);
        if (o != 0) return o > 0;
        // p on the hull face's line / plane: it conflicts with the region beyond that face exactly
        // when it lies inside the face's circumcircle, which is when it conflicts with the finite
        // simplex on the other side (whose circumsphere cuts that plane in the same circle)
        return conflict_finite(S_[s].n[i], p);
    }

    int new_simplex()
    {
        if (!free_.empty()) {
            const int s = free_.back();
            free_.pop_back();
            return s;
        }
        S_.push_back(Simplex{});
        mark_.push_back(0);
        return (int)S_.size() - 1;
    }

    bool init_simplex(const std::vector<int>& order)
    {
        // the first D+1 affinely independent points of the insertion order
        int pick[4] = {-1, -1, -1, -1};
        int got = 0;
        for (int i : order) {
            if (got == 0) {
                pick[got++] = i;
            } else if (got == 1) {
                bool same = true;
                for (int k = 0; k < D; ++k) same = same && pt(i)[k] == pt(pick[0])[k];
                if (!same) pick[got++] = i;
            } else if (got == 2) {
                if constexpr (D == 2) {
                    if (orient2(pt(pick[0]), pt(pick[1]), pt(i)) != 0) pick[got++] = i;
                } else {
                    // not collinear: some coordinate-plane projection has a nonzero orientation
                    double a[2], b[2], c[2];
                    bool ok = false;
                    for (int pr = 0; pr < 3 && !ok; ++pr) {
                        const int x = pr == 2 ? 1 : 0, y = pr == 0 ? 1 : 2;
                        a[0] = pt(pick[0])[x]; a[1] = pt(pick[0])[y];
                        b[0] = pt(pick[1])[x]; b[1] = pt(pick[1])[y];
                        c[0] = pt(i)[x]; c[1] = pt(i)[y];
                        ok = orient2(a, b, c) != 0;
                    }
                    if (ok) pick[got++] = i;
                }
            } else if (D == 3 && got == 3) {
                if (orient3(pt(pick[0]), pt(pick[1]), pt(pick[2]), pt(i)) != 0) pick[got++] = i;
            }
            if (got == NV) break;
        }
        if (got < NV) return false;
        for (int k = 0; k < NV; ++k) init_[k] = pick[k];
        int v[NV];
        for (int k = 0; k < NV; ++k) v[k] = pick[k];
        int o;
        if constexpr (D == 2)
            o = orient2(pt(v[0]), pt(v[1]), pt(v[2]));
        else
            o = orient3(pt(v[0]), pt(v[1]), pt(v[2]), pt(v[3]));
        if (o < 0) std::swap(v[0], v[1]);
        const int s0 = new_simplex();
        for (int k = 0; k < NV; ++k) S_[s0].v[k] = v[k];
        // one infinite simplex beyond every face: the face reversed, the infinite vertex last
        int inf_s[NV];
        for (int i = 0; i < NV; ++i) {
            int f[D];
            face(s0, i, f);
            std::swap(f[0], f[1]);
            const int s = new_simplex();
            for (int k = 0; k < D; ++k) S_[s].v[k] = f[k];
            S_[s].v[D] = inf_;
            S_[s].n[D] = s0;
            S_[s0].n[i] = s;
            inf_s[i] = s;
        }
        link_by_ridges(inf_s, NV);
        last_ = s0;
        return true;
    }

    // link the faces of the given simplices that hold their last vertex (slot D): every such
    // face is shared by two of them.  The face opposite slot j < D holds the other D-1 vertices
    // of the first D slots (its ridge): 2D the vertex at slot 1-j, 3D the edge of the other two.
    void link_by_ridges(const int* simp, int count)
    {
        size_t cap = 256;
        while (cap < (size_t)count * D * 2) cap <<= 1;
        if (ht_.size() < cap) {
            ht_.assign(cap, HEntry{0, 0, 0});
            gen_ = 0;
        }
        if (++gen_ == 0) {
            for (auto& e : ht_) e.gen = 0;
            gen_ = 1;
        }
        const size_t mask = ht_.size() - 1;
        for (int c = 0; c < count; ++c) {
            const int s = simp[c];
            const int* v = S_[s].v;
            for (int j = 0; j < D; ++j) {
                uint64_t key;
                if constexpr (D == 2) {
                    key = (uint32_t)v[1 - j];
                } else {
                    const int a = v[(j + 1) % 3], b = v[(j + 2) % 3];
                    key = a < b ? ((uint64_t)(uint32_t)a << 32) | (uint32_t)b : ((uint64_t)(uint32_t)b << 32) | (uint32_t)a;
                }
                size_t h = (size_t)((key * 0x9E3779B97F4A7C15ULL) >> 32) & mask;
                while (ht_[h].gen == gen_ && ht_[h].key != key) h = (h + 1) & mask;
                HEntry& e = ht_[h];
                if (e.gen == gen_) {
                    const int t = e.val / NV, tj = e.val % NV;
                    S_[s].n[j] = t;
                    S_[t].n[tj] = s;
                    e.key = ~0ULL;  // matched: stays occupied (probe chains intact), equals no key
                } else {
                    e.key = key;
                    e.val = s * NV + j;
                    e.gen = gen_;
                }
            }
        }
    }

    int locate(const double* p)
    {
        int s = last_;
        if (!alive(s)) s = 0;
        while (!alive(s)) ++s;
        {
            const int i = inf_slot(s);
            if (i >= 0) s = S_[s].n[i];
        }
        int prev = -1;
        for (size_t guard = 0; guard < 4 * S_.size() + 64; ++guard) {
            if (inf_slot(s) >= 0) return s;
            const int r = (int)(rng_() % NV);
            int next = -1;
            for (int q = 0; q < NV; ++q) {
                const int i = (r + q) % NV;
                if (S_[s].n[i] == prev) continue;
                int f[D];
                face(s, i, f);
                if (orient(f, p) < 0) {
                    next = S_[s].n[i];
                    break;
                }
            }
            if (next < 0) {
                // also the face we came through: p may lie beyond it only through rounding-free
                // ties, which the exact predicates exclude
                return s;
            }
            prev = s;
            s = next;
        }
        throw std::runtime_error("point location did not terminate");
    }

    bool insert(int id)
    {
        const double* p = pt(id);
        const int s0 = locate(p);
        if (inf_slot(s0) < 0) {
            for (int k = 0; k < NV; ++k) {
                const double* q = pt(S_[s0].v[k]);
                bool same = true;
                for (int c = 0; c < D; ++c) same = same && q[c] == p[c];
                if (same) return false;  // an exact duplicate of a vertex: not inserted
            }
        }
        ++stamp_;
        if (stamp_ == 0) {
            std::fill(mark_.begin(), mark_.end(), 0);
            stamp_ = 1;
        }
        // cavity: the simplices in conflict with p (a ball around p); marks: stamp = in the cavity
        cavity_.clear();
        bfaces_.clear();
        cavity_.push_back(s0);
        mark_[s0] = stamp_;
        for (size_t c = 0; c < cavity_.size(); ++c) {
            const int s = cavity_[c];
            for (int i = 0; i < NV; ++i) {
                const int t = S_[s].n[i];
                if (mark_[t] == stamp_) continue;
                if (conflict(t, p)) {
                    mark_[t] = stamp_;
                    cavity_.push_back(t);
                } else {
                    bfaces_.push_back(s * NV + i);
                }
            }
        }
        // a simplex tested out of the cavity from two sides is listed once per face: fine
        created_.clear();
        for (int bf : bfaces_) {
            const int s = bf / NV, i = bf % NV;
            const int out = S_[s].n[i];
            int f[D];
            face(s, i, f);
            const int ns = new_simplex();
            for (int k = 0; k < D; ++k) S_[ns].v[k] = f[k];
            S_[ns].v[D] = id;
            S_[ns].n[D] = out;
            for (int k = 0; k < NV; ++k)
                if (S_[out].n[k] == s) S_[out].n[k] = ns;
            created_.push_back(ns);
        }
        link_by_ridges(created_.data(), (int)created_.size());
        for (int s : cavity_) {
            S_[s].v[0] = -1;
            free_.push_back(s);
        }
        // new slots may have been taken from the free list before this insertion's frees: make
        // sure the marks of recycled slots cannot alias this stamp
        for (int s : created_) mark_[s] = 0;
        last_ = created_.empty() ? last_ : created_[0];
        return true;
    }
};

// ---------------------------------------------------------------------------------------------
// geometry helpers
static double det_j(int dim, const double* P, const int* v)  // Element::computeJ + computeDetJ (Element.cpp:15-86)
{
    if (dim == 2) {
        const double* a = P + 2 * v[0];
        const double* b = P + 2 * v[1];
        const double* c = P + 2 * v[2];
        const double J00 = b[0] - a[0], J01 = c[0] - a[0], J10 = b[1] - a[1], J11 = c[1] - a[1];
        return J00 * J11 - J10 * J01;
    }
    const double* a = P + 3 * v[0];
    const double* b = P + 3 * v[1];
    const double* c = P + 3 * v[2];
    const double* d = P + 3 * v[3];
    double J[3][3];
    for (int k = 0; k < 3; ++k) {
        J[k][0] = b[k] - a[k];
        J[k][1] = c[k] - a[k];
        J[k][2] = d[k] - a[k];
    }
    return J[0][0] * J[1][1] * J[2][2] + J[0][1] * J[1][2] * J[2][0] + J[0][2] * J[1][0] * J[2][1] -
           J[2][0] * J[1][1] * J[0][2] - J[2][1] * J[1][2] * J[0][0] - J[2][2] * J[1][0] * J[0][1];
}

static double circumradius2(int dim, const double* P, const int* v)
{
    if (dim == 2) {
        const double* a = P + 2 * v[0];
        const double bx = P[2 * v[1]] - a[0], by = P[2 * v[1] + 1] - a[1];
        const double cx = P[2 * v[2]] - a[0], cy = P[2 * v[2] + 1] - a[1];
        const double b2 = bx * bx + by * by, c2 = cx * cx + cy * cy;
        const double den = 2.0 * (bx * cy - by * cx);
        const double ux = (cy * b2 - by * c2) / den, uy = (bx * c2 - cx * b2) / den;
        return ux * ux + uy * uy;
    }
    const double* a = P + 3 * v[0];
    double e[3][3], l[3];
    for (int i = 0; i < 3; ++i) {
        for (int k = 0; k < 3; ++k) e[i][k] = P[3 * v[i + 1] + k] - a[k];
        l[i] = e[i][0] * e[i][0] + e[i][1] * e[i][1] + e[i][2] * e[i][2];
    }
    auto cross = [](const double* x, const double* y, double* z) {
        z[0] = x[1] * y[2] - x[2] * y[1];
        z[1] = x[2] * y[0] - x[0] * y[2];
        z[2] = x[0] * y[1] - x[1] * y[0];
    };
    double c01[3], c12[3], c20[3];
    cross(e[1], e[2], c12);
    cross(e[2], e[0], c20);
    cross(e[0], e[1], c01);
    const double det = e[0][0] * c12[0] + e[0][1] * c12[1] + e[0][2] * c12[2];
    double u[3];
    for (int k = 0; k < 3; ++k) u[k] = (l[0] * c12[k] + l[1] * c20[k] + l[2] * c01[k]) / (2.0 * det);
    return u[0] * u[0] + u[1] * u[1] + u[2] * u[2];
}

// canonical node tuple: the smallest node first, orientation kept (a cyclic rotation in 2D, an
// even permutation in 3D)
static void canonical(int dim, int* v)
{
    if (dim == 2) {
        int best[3] = {v[0], v[1], v[2]};
        for (int r = 1; r < 3; ++r) {
            const int c[3] = {v[r], v[(r + 1) % 3], v[(r + 2) % 3]};
            if (std::lexicographical_compare(c, c + 3, best, best + 3)) std::copy(c, c + 3, best);
        }
        std::copy(best, best + 3, v);
        return;
    }
    // the even permutation that brings the smallest node to slot 0, then the cyclic rotation of
    // the other three that brings their smallest to slot 1
    static const int FIRST[4][4] = {{0, 1, 2, 3}, {1, 0, 3, 2}, {2, 0, 1, 3}, {3, 0, 2, 1}};
    int k = 0;
    for (int j = 1; j < 4; ++j)
        if (v[j] < v[k]) k = j;
    const int a = v[FIRST[k][1]], b = v[FIRST[k][2]], c = v[FIRST[k][3]];
    int best[4];
    best[0] = v[k];
    if (a < b && a < c) {
        best[1] = a; best[2] = b; best[3] = c;
    } else if (b < c) {
        best[1] = b; best[2] = c; best[3] = a;
    } else {
        best[1] = c; best[2] = a; best[3] = b;
    }
    std::copy(best, best + 4, v);
}

// ---------------------------------------------------------------------------------------------
Mesher::Mesher(const MesherParams& prm) : prm_(prm)
{
    if (prm.dim != 2 && prm.dim != 3) throw std::invalid_argument("dim must be 2 or 3");
    if (!(prm.hchar > 0)) throw std::invalid_argument("hchar must be > 0");
    if (!(prm.alpha > 0)) throw std::invalid_argument("alpha must be > 0");
}

void Mesher::set_nodes(int n, int nstates, const double* pos_fb, const double* states_fb, const uint8_t* flags,
                       const int32_t* tags)
{
    const int d = prm_.dim;
    if (n <= d) throw std::invalid_argument("too few nodes");
    if (nstates < 0) throw std::invalid_argument("nstates < 0");
    N_ = n;
    ns_ = nstates;
    pos_.assign((size_t)n * d, 0.0);
    for (int i = 0; i < n; ++i)
        for (int k = 0; k < d; ++k) pos_[(size_t)d * i + k] = pos_fb[(size_t)k * n + i];
    st_.assign((size_t)n * nstates, 0.0);
    for (int i = 0; i < n; ++i)
        for (int s = 0; s < nstates; ++s) st_[(size_t)nstates * i + s] = states_fb ? states_fb[(size_t)s * n + i] : 0.0;
    flags_.assign(n, 0);
    for (int i = 0; i < n; ++i) flags_[i] = flags ? (uint8_t)(flags[i] & (kBound | kFixed | kDirichlet | kDirichletT)) : 0;
    tag_.assign(n, -1);
    if (tags)
        for (int i = 0; i < n; ++i) tag_[i] = tags[i];
    elems_.clear();
    facets_.clear();
    fout_.clear();
    felem_.clear();
    nbr_ptr_.assign(n + 1, 0);
    nbr_.clear();
    nelem_.assign(n, 0);
    triangulated_ = false;
}

void Mesher::set_positions_states(const double* pos_fb, const double* states_fb)
{
    const int d = prm_.dim, n = N_;
    if (pos_fb)
        for (int i = 0; i < n; ++i)
            for (int k = 0; k < d; ++k) pos_[(size_t)d * i + k] = pos_fb[(size_t)k * n + i];
    if (states_fb)
        for (int i = 0; i < n; ++i)
            for (int s = 0; s < ns_; ++s) st_[(size_t)ns_ * i + s] = states_fb[(size_t)s * n + i];
}

bool Mesher::is_onfs(int i) const { return (flags_[i] & kOnFS) != 0; }
bool Mesher::is_bound(int i) const { return (flags_[i] & kBound) != 0; }

// Mesh::triangulateAlphaShape2D / 3D (Mesh2D.cpp:32-251, Mesh3D.cpp:36-286)
void Mesher::triangulate()
{
    using clk = std::chrono::steady_clock;
    const auto t0 = clk::now();
    const int d = prm_.dim, nv = d + 1, n = N_;
    for (int i = 0; i < n; ++i) flags_[i] &= (uint8_t) ~(kOnFS | kFree);
    // 1. Delaunay triangulation; simplices in a flat list, neighbours across faces (-1: hull)
    std::vector<int> simp, nb;
    int ns = 0;
    std::vector<int> skipped;
    auto collect = [&](auto& tri) {
        using T = std::decay_t<decltype(tri)>;
        const auto& S = tri.simplices();
        std::vector<int> idx(S.size(), -1);
        for (size_t s = 0; s < S.size(); ++s)
            if (tri.alive((int)s) && !tri.infinite((int)s)) idx[s] = ns++;
        simp.resize((size_t)ns * nv);
        nb.resize((size_t)ns * nv);
        for (size_t s = 0; s < S.size(); ++s) {
            if (idx[s] < 0) continue;
            for (int k = 0; k < T::NV; ++k) {
                simp[(size_t)idx[s] * nv + k] = S[s].v[k];
                nb[(size_t)idx[s] * nv + k] = idx[S[s].n[k]];
            }
        }
    };
    if (d == 2) {
        Triangulation<2> tri;
        skipped = tri.build(pos_.data(), n);
        collect(tri);
    } else {
        Triangulation<3> tri;
        skipped = tri.build(pos_.data(), n);
        collect(tri);
    }
    const auto t1 = clk::now();
    // 2. alpha classification: interior iff squared circumradius <= alpha^2 hchar^2
    const double alpha2 = prm_.alpha * prm_.alpha * prm_.hchar * prm_.hchar;
    std::vector<uint8_t> interior(ns), keep(ns, 0);
    for (int s = 0; s < ns; ++s) interior[s] = circumradius2(d, pos_.data(), &simp[(size_t)s * nv]) <= alpha2;
    // vertex -> incident finite simplices (for the all-bound deletion rule)
    std::vector<int> vptr(n + 1, 0), vs;
    for (int s = 0; s < ns; ++s)
        for (int k = 0; k < nv; ++k) {
            const int v = simp[(size_t)s * nv + k];
            if (is_bound(v)) vptr[v + 1]++;
        }
    for (int i = 0; i < n; ++i) vptr[i + 1] += vptr[i];
    vs.resize(vptr[n]);
    {
        std::vector<int> fill(vptr.begin(), vptr.end() - 1);
        for (int s = 0; s < ns; ++s)
            for (int k = 0; k < nv; ++k) {
                const int v = simp[(size_t)s * nv + k];
                if (is_bound(v)) vs[fill[v]++] = s;
            }
    }
    const double h3 = prm_.hchar * prm_.hchar * prm_.hchar;
    // checkFaceDeletion / checkCellDeletion: true = delete
    auto deleted = [&](int s) -> bool {
        const int* v = &simp[(size_t)s * nv];
        double mean[3] = {0, 0, 0};
        if (!prm_.exclusion_zones.empty())
            for (int k = 0; k < d; ++k) {
                for (int j = 0; j < nv; ++j) mean[k] += pos_[(size_t)d * v[j] + k];
                mean[k] /= nv;
            }
        for (const auto& z : prm_.exclusion_zones) {
            bool in = true;
            for (int k = 0; k < d; ++k) in = in && mean[k] > z[k] && mean[k] < z[d + k];
            if (in) return true;
        }
        for (int j = 0; j < nv; ++j)
            if (!is_bound(v[j])) return false;
        if (d == 3 && det_j(3, pos_.data(), v) / 6.0 < 0.02 * h3) return true;
        // keep it when some alpha-interior simplex around one of its nodes reaches a non-boundary
        // node outside it
        for (int j = 0; j < nv; ++j) {
            for (int q = vptr[v[j]]; q < vptr[v[j] + 1]; ++q) {
                const int t = vs[q];
                if (!interior[t]) continue;
                for (int k = 0; k < nv; ++k) {
                    const int w = simp[(size_t)t * nv + k];
                    bool own = false;
                    for (int m = 0; m < nv; ++m) own = own || w == v[m];
                    if (!own && !is_bound(w)) return false;
                }
            }
        }
        return true;
    };
    for (int s = 0; s < ns; ++s) keep[s] = interior[s] && !deleted(s);
    // 3. kept elements in canonical form and order
    std::vector<std::pair<std::array<int, 4>, int>> kept;
    {
        // bucket by the first (smallest) node, then sort each bucket: sorted tuples overall
        std::vector<int> cnt(n + 1, 0);
        std::vector<std::array<int, 4>> canon(ns);
        int nk = 0;
        for (int s = 0; s < ns; ++s) {
            if (!keep[s]) continue;
            std::array<int, 4> v = {-1, -1, -1, -1};
            for (int k = 0; k < nv; ++k) v[k] = simp[(size_t)s * nv + k];
            canonical(d, v.data());
            canon[s] = v;
            cnt[v[0] + 1]++;
            ++nk;
        }
        if (nk == 0)
            throw std::runtime_error("Something went wrong while remeshing. You might have not chosen a good \"hchar\" "
                                     "with regard to your .msh file");
        for (int i = 0; i < n; ++i) cnt[i + 1] += cnt[i];
        kept.resize(nk);
        for (int s = 0; s < ns; ++s)
            if (keep[s]) kept[cnt[canon[s][0]]++] = {canon[s], s};
        for (int i = n; i > 0; --i) cnt[i] = cnt[i - 1];
        cnt[0] = 0;
        for (int i = 0; i < n; ++i)
            if (cnt[i + 1] - cnt[i] > 1) std::sort(kept.begin() + cnt[i], kept.begin() + cnt[i + 1]);
    }
    std::vector<int> elem_of(ns, -1);
    elems_.resize(kept.size() * nv);
    for (size_t e = 0; e < kept.size(); ++e) {
        for (int k = 0; k < nv; ++k) elems_[e * nv + k] = kept[e].first[k];
        elem_of[kept[e].second] = (int)e;
    }
    // 4. free-surface facets: faces of a kept interior simplex whose other side is not interior
    facets_.clear();
    fout_.clear();
    felem_.clear();
    std::vector<std::pair<std::array<int, 5>, int>> fl;  // (canonical nodes, out, elem), index
    for (int s = 0; s < ns; ++s) {
        if (!keep[s]) continue;
        for (int i = 0; i < nv; ++i) {
            const int t = nb[(size_t)s * nv + i];
            if (t >= 0 && interior[t]) continue;
            std::array<int, 5> f = {-1, -1, -1, -1, -1};
            if (d == 2) {
                for (int k = 0; k < 2; ++k) f[k] = simp[(size_t)s * nv + Face<2>::T[i][k]];
            } else {
                int tri[3];
                for (int k = 0; k < 3; ++k) tri[k] = simp[(size_t)s * nv + Face<3>::T[i][k]];
                // cyclic rotation: smallest node first
                const int r = tri[0] < tri[1] ? (tri[0] < tri[2] ? 0 : 2) : (tri[1] < tri[2] ? 1 : 2);
                for (int k = 0; k < 3; ++k) f[k] = tri[(r + k) % 3];
            }
            f[3] = simp[(size_t)s * nv + i];
            f[4] = elem_of[s];
            fl.push_back({f, 0});
        }
    }
    std::sort(fl.begin(), fl.end());
    for (const auto& it : fl) {
        for (int k = 0; k < d; ++k) facets_.push_back(it.first[k]);
        fout_.push_back(it.first[3]);
        felem_.push_back(it.first[4]);
        bool all_bound = true;
        for (int k = 0; k < d; ++k) all_bound = all_bound && is_bound(it.first[k]);
        if (!all_bound)
            for (int k = 0; k < d; ++k) flags_[it.first[k]] |= kOnFS;
    }
    // 5. node -> neighbours (sorted, unique) and element counts
    build_adjacency();
    const auto t2 = clk::now();
    skipped_ = (int)skipped.size();
    t_delaunay_ = std::chrono::duration<double>(t1 - t0).count();
    t_alpha_ = std::chrono::duration<double>(t2 - t1).count();
    triangulated_ = true;
    if (prm_.delete_flying_nodes) delete_flying_nodes();
}

void Mesher::build_adjacency()
{
    const int d = prm_.dim, nv = d + 1, n = N_;
    const int E = (int)(elems_.size() / nv);
    nelem_.assign(n, 0);
    for (int v : elems_) nelem_[v]++;
    // node -> elements, then each node's neighbours deduplicated with a stamp and sorted
    std::vector<int> eptr(n + 1, 0);
    for (int i = 0; i < n; ++i) eptr[i + 1] = eptr[i] + nelem_[i];
    std::vector<int> eidx(eptr[n]), fill(eptr.begin(), eptr.end() - 1);
    for (int e = 0; e < E; ++e)
        for (int a = 0; a < nv; ++a) eidx[fill[elems_[(size_t)e * nv + a]]++] = e;
    std::vector<int> seen(n, -1);
    nbr_ptr_.assign(n + 1, 0);
    nbr_.clear();
    nbr_.reserve((size_t)eptr[n] * 2);
    for (int i = 0; i < n; ++i) {
        const size_t b = nbr_.size();
        for (int q = eptr[i]; q < eptr[i + 1]; ++q) {
            const int* v = &elems_[(size_t)eidx[q] * nv];
            for (int a = 0; a < nv; ++a)
                if (v[a] != i && seen[v[a]] != i) {
                    seen[v[a]] = i;
                    nbr_.push_back(v[a]);
                }
        }
        std::sort(nbr_.begin() + b, nbr_.end());
        nbr_ptr_[i + 1] = (int)nbr_.size();
    }
    for (int i = 0; i < n; ++i)
        if (nelem_[i] == 0) flags_[i] |= kFree;
}

// drops the nodes flagged in `del` (stable); elements, facets and neighbour lists are renumbered
// when `renumber` (the reference's deleteFlyingNodes renumbers elements and neighbour lists only,
// Mesh.cpp:308-330, leaving facet indices stale; they are renumbered here as well)
void Mesher::erase_nodes(const std::vector<uint8_t>& del, bool renumber)
{
    const int d = prm_.dim, n = N_;
    std::vector<int> newid(n, -1);
    int m = 0;
    for (int i = 0; i < n; ++i)
        if (!del[i]) newid[i] = m++;
    if (m == n) return;
    std::vector<double> pos((size_t)m * d), st((size_t)m * ns_);
    std::vector<uint8_t> fl(m);
    std::vector<int32_t> tg(m);
    std::vector<int> ne(m);
    for (int i = 0; i < n; ++i) {
        const int j = newid[i];
        if (j < 0) continue;
        for (int k = 0; k < d; ++k) pos[(size_t)d * j + k] = pos_[(size_t)d * i + k];
        for (int s = 0; s < ns_; ++s) st[(size_t)ns_ * j + s] = st_[(size_t)ns_ * i + s];
        fl[j] = flags_[i];
        tg[j] = tag_[i];
        ne[j] = nelem_.empty() ? 0 : nelem_[i];
    }
    pos_.swap(pos);
    st_.swap(st);
    flags_.swap(fl);
    tag_.swap(tg);
    nelem_.swap(ne);
    if (renumber) {
        for (int& v : elems_) v = newid[v];
        for (int& v : facets_) v = newid[v];
        for (int& v : fout_) v = newid[v];
        std::vector<int> ptr(m + 1, 0), nb;
        for (int i = 0; i < n; ++i) {
            const int j = newid[i];
            if (j < 0) continue;
            for (int q = nbr_ptr_[i]; q < nbr_ptr_[i + 1]; ++q)
                if (newid[nbr_[q]] >= 0) nb.push_back(newid[nbr_[q]]);
            ptr[j + 1] = (int)nb.size();
        }
        nbr_ptr_.swap(ptr);
        nbr_.swap(nb);
    } else {
        // node lists only: adjacency is stale until the next triangulation
        nbr_ptr_.assign(m + 1, 0);
        nbr_.clear();
    }
    N_ = m;
}

// Mesh::deleteFlyingNodes (Mesh.cpp:251-331): nodes in no element that are not boundary nodes
void Mesher::delete_flying_nodes()
{
    std::vector<uint8_t> del(N_, 0);
    int cnt = 0;
    for (int i = 0; i < N_; ++i)
        if (nelem_[i] == 0 && !is_bound(i)) {
            del[i] = 1;
            ++cnt;
        }
    flying_deleted_ = cnt;
    if (cnt) erase_nodes(del, true);
}

// Mesh::laplacianSmoothingBoundaries (Mesh.cpp:580-760)
void Mesher::laplacian_smoothing_boundaries()
{
    if (!prm_.laplacian_smoothing_boundaries) return;
    const int d = prm_.dim, nv = d + 1;
    const double h = prm_.hchar;
    const int E = (int)(elems_.size() / nv);
    for (int e = 0; e < E; ++e) {
        const int* v = &elems_[(size_t)e * nv];
        int out = -1, fn[3] = {-1, -1, -1}, cnt = 0;
        bool contact = false;
        for (int k = 0; k < nv; ++k) contact = contact || is_bound(v[k]);
        if (!contact) continue;
        for (int k = 0; k < nv; ++k) {
            if (!is_bound(v[k]))
                out = v[k];
            else
                fn[cnt++] = v[k];
        }
        if (cnt != d) continue;
        double* po = &pos_[(size_t)d * out];
        auto P = [&](int node, int k) { return pos_[(size_t)d * node + k]; };
        double nrm[3] = {0, 0, 0};
        double dist = 0;
        if (d == 2) {
            if (P(fn[0], 0) == P(fn[1], 0)) {
                nrm[0] = 1;
                dist = std::fabs(P(fn[0], 0) - po[0]);
            } else if (P(fn[0], 1) == P(fn[1], 1)) {
                nrm[1] = 1;
                dist = std::fabs(P(fn[0], 1) - po[1]);
            } else {
                // the line a x + b y + 1 = 0 through both facet nodes
                const double x0 = P(fn[0], 0), y0 = P(fn[0], 1), x1 = P(fn[1], 0), y1 = P(fn[1], 1);
                const double det = x0 * y1 - x1 * y0;
                if (det == 0) continue;  // a line through the origin (see the 3D case)
                nrm[0] = (-y1 + y0) / det;
                nrm[1] = (-x0 + x1) / det;
                dist = std::fabs(nrm[0] * po[0] + nrm[1] * po[1] + 1) / std::sqrt(nrm[0] * nrm[0] + nrm[1] * nrm[1]);
            }
        } else {
            if (P(fn[0], 0) == P(fn[1], 0) && P(fn[1], 0) == P(fn[2], 0)) {
                nrm[0] = 1;
                dist = std::fabs(P(fn[0], 0) - po[0]);
            } else if (P(fn[0], 1) == P(fn[1], 1) && P(fn[1], 1) == P(fn[2], 1)) {
                nrm[1] = 1;
                dist = std::fabs(P(fn[0], 1) - po[1]);
            } else if (P(fn[0], 2) == P(fn[1], 2) && P(fn[1], 2) == P(fn[2], 2)) {
                // (the reference compares the second node's y with the third node's z here,
                // Mesh.cpp:662, which sends facets of a z = const wall to the plane solve below;
                // for a wall through the origin that system is singular.  Corrected here.)
                nrm[2] = 1;
                dist = std::fabs(P(fn[0], 2) - po[2]);
            } else {
                // the plane a x + b y + c z + 1 = 0 through the three facet nodes (Cramer's rule)
                double A[3][3];
                for (int r = 0; r < 3; ++r)
                    for (int k = 0; k < 3; ++k) A[r][k] = P(fn[r], k);
                auto det3 = [](double M[3][3]) {
                    return M[0][0] * (M[1][1] * M[2][2] - M[1][2] * M[2][1]) -
                           M[0][1] * (M[1][0] * M[2][2] - M[1][2] * M[2][0]) +
                           M[0][2] * (M[1][0] * M[2][1] - M[1][1] * M[2][0]);
                };
                const double D0 = det3(A);
                // a plane through the origin has no such equation (the reference's least-squares
                // QR solve returns an arbitrary normal): the element is left alone
                if (D0 == 0) continue;
                for (int k = 0; k < 3; ++k) {
                    double B[3][3];
                    std::memcpy(B, A, sizeof(B));
                    for (int r = 0; r < 3; ++r) B[r][k] = -1.0;
                    nrm[k] = det3(B) / D0;
                }
                dist = std::fabs(nrm[0] * po[0] + nrm[1] * po[1] + nrm[2] * po[2] + 1) /
                       std::sqrt(nrm[0] * nrm[0] + nrm[1] * nrm[1] + nrm[2] * nrm[2]);
            }
        }
        if (dist > h / 3) continue;
        if (is_onfs(out)) {
            double toward[3] = {0, 0, 0};
            for (int k = 0; k < d; ++k) toward[k] = po[k] - P(fn[0], k);
            if (d == 2) {
                if (nrm[0] * toward[0] + nrm[1] * toward[1] < 0) {
                    nrm[0] = -nrm[0];
                    nrm[1] = -nrm[1];
                }
                const double nn = std::sqrt(nrm[0] * nrm[0] + nrm[1] * nrm[1]);
                nrm[0] /= nn;
                nrm[1] /= nn;
            } else {
                const double nn = std::sqrt(nrm[0] * nrm[0] + nrm[1] * nrm[1] + nrm[2] * nrm[2]);
                for (int k = 0; k < 3; ++k) nrm[k] /= nn;
                if (toward[0] * nrm[0] + toward[1] * nrm[1] + toward[2] * nrm[2] < 0)
                    for (int k = 0; k < 3; ++k) nrm[k] = -nrm[k];
            }
            for (int k = 0; k < d; ++k) po[k] += h / 3 * nrm[k];
        } else {
            const int b = nbr_ptr_[out], c = nbr_ptr_[out + 1];
            double np[3] = {0, 0, 0};
            for (int q = b; q < c; ++q)
                for (int k = 0; k < d; ++k) np[k] += P(nbr_[q], k);
            for (int k = 0; k < d; ++k) po[k] = np[k] / static_cast<double>(c - b);
        }
    }
}

// Mesh::addNodes (Mesh.cpp:27-147): a node at the centre of every element larger than
// omega hchar^dim (not on the free surface unless addOnFS), states averaged
int Mesher::add_nodes()
{
    const int d = prm_.dim, nv = d + 1;
    const int E = (int)(elems_.size() / nv);
    const double limit = prm_.omega * std::pow(prm_.hchar, d);
    const double ref = d == 2 ? 0.5 : 1.0 / 6.0;
    int added = 0;
    for (int e = 0; e < E; ++e) {
        const int* v = &elems_[(size_t)e * nv];
        const double size = det_j(d, pos_.data(), v) * ref;
        bool on_fs = false;
        for (int k = 0; k < nv; ++k) on_fs = on_fs || is_onfs(v[k]);
        if (!(size > limit && (prm_.add_on_fs || !on_fs))) continue;
        for (int k = 0; k < d; ++k) {
            double s = 0;
            for (int j = 0; j < nv; ++j) s += pos_[(size_t)d * v[j] + k];
            pos_.push_back(s / nv);
        }
        for (int q = 0; q < ns_; ++q) {
            double s = 0;
            for (int j = 0; j < nv; ++j) s += st_[(size_t)ns_ * v[j] + q];
            st_.push_back(s / nv);
        }
        flags_.push_back(kFree);  // in no element until the next triangulation
        tag_.push_back(-1);
        nelem_.push_back(0);
        nbr_ptr_.push_back(nbr_ptr_.back());
        ++N_;
        ++added;
    }
    return added;
}

// Mesh::removeNodes (Mesh.cpp:928-1017): of two nodes closer than gamma hchar (half that between
// two free-surface nodes) one is deleted, never a boundary node nor a free-surface node next to
// an interior one
int Mesher::remove_nodes()
{
    const int d = prm_.dim, n = N_;
    std::vector<uint8_t> touched(n, 0), del(n, 0);
    const double limit = prm_.gamma * prm_.hchar;
    int removed = 0;
    for (int i = 0; i < n; ++i) {
        if (del[i] || nelem_[i] == 0) continue;
        for (int q = nbr_ptr_[i]; q < nbr_ptr_[i + 1]; ++q) {
            const int j = nbr_[q];
            double dd = 0;
            for (int k = 0; k < d; ++k) {
                const double t = pos_[(size_t)d * i + k] - pos_[(size_t)d * j + k];
                dd += t * t;
            }
            dd = std::sqrt(dd);
            const double fact = (is_onfs(i) && is_onfs(j)) ? 0.5 : 1.0;
            if (!(dd <= fact * limit)) continue;
            if (touched[j]) {
                if (is_bound(i) || (is_onfs(i) && !is_onfs(j))) continue;
                if (!del[i]) ++removed;
                del[i] = 1;
            } else {
                if (is_bound(j) || (is_onfs(j) && !is_onfs(i))) continue;
                touched[i] = 1;
                if (!del[j]) ++removed;
                del[j] = 1;
            }
        }
    }
    if (removed) erase_nodes(del, false);
    return removed;
}

// Mesh::checkBoundingBox (Mesh.cpp:149-207)
int Mesher::check_bounding_box()
{
    if (!prm_.has_bounding_box) return 0;
    const int d = prm_.dim, n = N_;
    std::vector<uint8_t> del(n, 0);
    int cnt = 0;
    for (int i = 0; i < n; ++i) {
        for (int k = 0; k < d; ++k) {
            const double x = pos_[(size_t)d * i + k];
            if (x < prm_.bounding_box[k] || x > prm_.bounding_box[d + k]) {
                del[i] = 1;
                break;
            }
        }
        cnt += del[i];
    }
    if (cnt) erase_nodes(del, false);
    return cnt;
}

// Mesh::remesh (Mesh.cpp:919-926)
void Mesher::remesh()
{
    if (!triangulated_) throw std::logic_error("triangulate the mesh before remeshing it");
    using clk = std::chrono::steady_clock;
    const auto t0 = clk::now();
    laplacian_smoothing_boundaries();
    added_ = add_nodes();
    removed_ = remove_nodes();
    out_of_box_ = check_bounding_box();
    const auto t1 = clk::now();
    t_nodes_ = std::chrono::duration<double>(t1 - t0).count();
    triangulate();
}

void Mesher::get_nodes(double* pos_fb, double* states_fb, uint8_t* flags, int32_t* tags) const
{
    const int d = prm_.dim, n = N_;
    if (pos_fb)
        for (int i = 0; i < n; ++i)
            for (int k = 0; k < d; ++k) pos_fb[(size_t)k * n + i] = pos_[(size_t)d * i + k];
    if (states_fb)
        for (int i = 0; i < n; ++i)
            for (int s = 0; s < ns_; ++s) states_fb[(size_t)s * n + i] = st_[(size_t)ns_ * i + s];
    if (flags) std::copy(flags_.begin(), flags_.end(), flags);
    if (tags) std::copy(tag_.begin(), tag_.end(), tags);
}

}  // namespace mesh
}  // namespace pfem
