// The heat equation of the reference (HeatEqIncompNewton without phase change,
// srcs/simulation/physics/IncompNewton/HeatEquation.inl), one conduction step on the GPU:
//
//   A = M + dt L on rows that are neither Dirichlet nor free (m_buildAb, :209-280), unit rows
//   otherwise; b = sum_e M_e theta_prev; then m_applyBC (:283-364): the facet terms
//   (prescribed flux q.n, convection h (T - Tinf), radiation sigma eps (T^4 - Tinf^4), each
//   integrated with the reference's 3-point facet rule) subtracted times dt in facet order,
//   free rows b = theta_prev, Dirichlet rows b = T_bc with the column elimination; and the
//   solve of the reference's Eigen::ConjugateGradient<Lower|Upper> (diagonal preconditioner,
//   initial guess theta_prev; HeatEquation.hpp:50, HeatEquation.inl:131).
//
// Layout: a scalar CSR over the case's node graph (a full row holds the sorted neighbours,
// the node itself included, so value k of row i is node-graph entry k of i; other rows hold
// their diagonal only).  The assembly gathers, like the PSPG one: each value sums the
// contributions of the elements holding both nodes (pair-element incidence, element order:
// M_e then dt L_e per element, the order of the reference's triplet list), from one element
// record per element (gradients and measure) written by an element pass.
#include "common.hpp"
#include "heat.hpp"
#include "p1geom.hpp"

namespace pfem {

namespace {

constexpr int kHeatRec = 16;  // doubles per element record: dN_i/dx_k (d (d+1)), |e| at 15

template <int D>
__global__ __launch_bounds__(kBlock) void k_heat_elem(HeatView V) {
    const int e = blockIdx.x * kBlock + threadIdx.x;
    if (e >= V.E) return;
    int nd[D + 1];
#pragma unroll
    for (int a = 0; a <= D; ++a) nd[a] = V.conn[(size_t)a * V.E + e];
    double x[D][D + 1];
    WGeom<D> G;
    wc_geometry<D>(V.pos, V.N, nd, x, G);
    double* rec = V.erec + (size_t)kHeatRec * e;
#pragma unroll
    for (int k = 0; k < D; ++k)
#pragma unroll
        for (int a = 0; a <= D; ++a) rec[k * (D + 1) + a] = G.gN[k][a];
    rec[15] = G.vol;
}

// one thread per node: its row of A = M + dt L (or the unit row) and, for a full row, its
// share of b = sum_e M_e theta_prev (the diagonal entry's pair-element list holds every element
// of the node, in element order)
template <int D>
__global__ __launch_bounds__(kBlock) void k_heat_asm(HeatView V) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= V.N) return;
    const uint8_t f = V.flags[i];
    const int rp = V.row_ptr[i];
    if ((f & V.dbit) || (f & NODE_FREE)) {
        V.vals[rp] = 1.0;
        return;
    }
    const double mfac = V.rho * V.cv / ((D + 1) * (D + 2));
    const int nb = V.nbr_ptr[i], deg = V.nbr_ptr[i + 1] - nb;
    for (int k = 0; k < deg; ++k) {
        const int g = nb + k;
        double acc = 0.0;
        for (int q = V.pei_ptr[g]; q < V.pei_ptr[g + 1]; ++q) {
            const unsigned pe = V.pei[q];
            const int e = (int)(pe >> 4), a = (int)((pe >> 2) & 3u), bb = (int)(pe & 3u);
            const double* rec = V.erec + (size_t)kHeatRec * e;
            const double vol = rec[15];
            double gg = 0;
#pragma unroll
            for (int c = 0; c < D; ++c) gg += rec[c * (D + 1) + a] * rec[c * (D + 1) + bb];
            acc += mfac * vol * (a == bb ? 2.0 : 1.0);
            acc += V.dt * (V.k * vol * gg);
        }
        V.vals[rp + k] = acc;
        if (V.nbr_idx[g] == i) {
            double bi = 0.0;
            for (int q = V.pei_ptr[g]; q < V.pei_ptr[g + 1]; ++q) {
                const unsigned pe = V.pei[q];
                const int e = (int)(pe >> 4), a = (int)((pe >> 2) & 3u);
                const double vol = V.erec[(size_t)kHeatRec * e + 15];
                double t = 0;
#pragma unroll
                for (int jj = 0; jj <= D; ++jj)
                    t += mfac * vol * (jj == a ? 2.0 : 1.0) * V.Tprev[V.conn[(size_t)jj * V.E + e]];
                bi += t;
            }
            V.b[i] = bi;
        }
    }
}

// one thread per facet: getQN (MatricesBuilder.inl:247-274) and the two getSGamma
// (:358-371) with the factors of HeatEquation.inl:87-99, T = N theta_prev at each point
template <int D>
__global__ __launch_bounds__(kBlock) void k_heat_facet(HeatView V) {
    const int f = blockIdx.x * kBlock + threadIdx.x;
    if (f >= V.F) return;
    int fn[D];
#pragma unroll
    for (int i = 0; i < D; ++i) fn[i] = V.facets[(size_t)f * D + i];
    const int out = V.fout[f];
    auto P = [&](int node, int k) { return V.pos[(size_t)k * V.N + node]; };
    double det, nrm[3] = {0, 0, 0};
    double Nq[3][D], w[3];
    double refs;
    if constexpr (D == 2) {
        const double J00 = (P(fn[1], 0) - P(fn[0], 0)) / 2, J10 = (P(fn[1], 1) - P(fn[0], 1)) / 2;
        det = sqrt(J00 * J00 + J10 * J10);
        double nx = P(fn[1], 1) - P(fn[0], 1), ny = P(fn[0], 0) - P(fn[1], 0);
        if (nx * (P(out, 0) - P(fn[0], 0)) + ny * (P(out, 1) - P(fn[0], 1)) > 0) {
            nx = -nx;
            ny = -ny;
        }
        const double nn = sqrt(nx * nx + ny * ny);
        nrm[0] = nx / nn;
        nrm[1] = ny / nn;
        const double xi[3] = {-sqrt(3.0 / 5.0), 0.0, sqrt(3.0 / 5.0)};
        w[0] = 5.0 / 18.0;
        w[1] = 8.0 / 18.0;
        w[2] = 5.0 / 18.0;
#pragma unroll
        for (int g = 0; g < 3; ++g) {
            Nq[g][0] = (1 - xi[g]) / 2;
            Nq[g][1] = (1 + xi[g]) / 2;
        }
        refs = 2.0;
    } else {
        double a[3], b[3], v[3];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            a[k] = P(fn[1], k) - P(fn[0], k);
            b[k] = P(fn[2], k) - P(fn[0], k);
            v[k] = P(out, k) - P(fn[0], k);
        }
        const double g0 = a[1] * b[2] - b[1] * a[2], g1 = a[2] * b[0] - b[2] * a[0], g2 = a[0] * b[1] - b[0] * a[1];
        det = sqrt(g0 * g0 + g1 * g1 + g2 * g2);
        double n[3] = {a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]};
        if (v[0] * n[0] + v[1] * n[1] + v[2] * n[2] > 0)
            for (int k = 0; k < 3; ++k) n[k] *= -1.0;
        const double nn = sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
        for (int k = 0; k < 3; ++k) nrm[k] = n[k] / nn;
        const double gp[3][2] = {{1.0 / 6.0, 1.0 / 6.0}, {1.0 / 6.0, 2.0 / 3.0}, {2.0 / 3.0, 1.0 / 6.0}};
#pragma unroll
        for (int g = 0; g < 3; ++g) {
            Nq[g][0] = 1 - gp[g][0] - gp[g][1];
            Nq[g][1] = gp[g][0];
            Nq[g][2] = gp[g][1];
            w[g] = 1.0 / 3.0;
        }
        refs = 0.5;
    }
    double qn[D], sgh[D], sgr[D];
#pragma unroll
    for (int i = 0; i < D; ++i) qn[i] = sgh[i] = sgr[i] = 0;
    const double sigma = 5.670374419e-8;
    const double T4inf = V.Tinf * V.Tinf * V.Tinf * V.Tinf;
#pragma unroll
    for (int g = 0; g < 3; ++g) {
        double qv = 0;
        if (V.qgp)
#pragma unroll
            for (int k = 0; k < D; ++k) qv += V.qgp[((size_t)f * 3 + g) * D + k] * nrm[k];
        const double fact = qv * w[g];
        double Tg = 0;
#pragma unroll
        for (int j = 0; j < D; ++j) Tg += Nq[g][j] * V.Tprev[fn[j]];
        const double fh = V.h * (Tg - V.Tinf);
        const double fr = sigma * V.epsRad * (Tg * Tg * Tg * Tg - T4inf);
#pragma unroll
        for (int i = 0; i < D; ++i) {
            qn[i] += fact * Nq[g][i];
            sgh[i] += fh * Nq[g][i] * w[g];
            sgr[i] += fr * Nq[g][i] * w[g];
        }
    }
    const double s = refs * det;
#pragma unroll
    for (int i = 0; i < D; ++i) {
        double* t = V.fterm + ((size_t)f * D + i) * 3;
        t[0] = qn[i] * s;
        t[1] = sgh[i] * s;
        t[2] = sgr[i] * s;
    }
}

// one thread per node: m_applyBC's effect on its row, in the reference's order (facet terms in
// facet order, then the eliminations of the Dirichlet columns in increasing node order), and
// the diagonal inverse for the preconditioner
__global__ __launch_bounds__(kBlock) void k_heat_bc(HeatView V) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= V.N) return;
    const uint8_t f = V.flags[i];
    const int rp = V.row_ptr[i];
    if (f & V.dbit) {
        V.b[i] = V.Tbc[i];
    } else if (f & NODE_FREE) {
        V.b[i] = V.Tprev[i];
    } else {
        double bi = V.b[i];
        for (int q = V.nf_ptr[i]; q < V.nf_ptr[i + 1]; ++q) {
            const int ent = V.nf_list[q];
            const int fc = ent / V.dim;
            const uint8_t kind = V.fkind[fc];
            const double* t = V.fterm + (size_t)ent * 3;
            if (kind & 1) bi -= V.dt * t[0];
            if (kind & 2) bi -= V.dt * t[1];
            if (kind & 4) bi -= V.dt * t[2];
        }
        const int nb = V.nbr_ptr[i], deg = V.nbr_ptr[i + 1] - nb;
        for (int k = 0; k < deg; ++k) {
            const int j = V.nbr_idx[nb + k];
            if (j != i && (V.flags[j] & V.dbit)) {
                bi -= V.vals[rp + k] * V.Tbc[j];
                V.vals[rp + k] = 0.0;
            }
        }
        V.b[i] = bi;
    }
    const double dg = V.vals[V.diag_pos[i]];
    V.dinv[i] = dg != 0.0 ? 1.0 / dg : 1.0;
}

// ---- conjugate gradient, diagonal preconditioner -------------------------------------------
// Reductions are fence-free: every workgroup writes its partial sums to its own slot, and the
// next kernel's workgroups re-reduce them in workgroup order.  rz partials alternate between
// two arrays by iteration parity, so no kernel reads a slot that it also writes.

__device__ double sum_partials(const double* p, int nb) {
    __shared__ double red[kBlock];
    double s = 0;
    for (int q = threadIdx.x; q < nb; q += kBlock) s += p[q];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = kBlock / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    const double r = red[0];
    __syncthreads();
    return r;
}

__device__ double row_dot(const HeatView& V, int i, const double* x) {
    double s = 0;
    for (int q = V.row_ptr[i]; q < V.row_ptr[i + 1]; ++q) s += V.vals[q] * x[V.col[q]];
    return s;
}

// r = b - A x, z = D^-1 r, p = z; partials of r.z (parity 0), r.r and b.b
__global__ __launch_bounds__(kBlock) void k_cg_init(HeatView V) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    double v[3] = {0, 0, 0};
    if (i < V.N) {
        const double r = V.b[i] - row_dot(V, i, V.x);
        const double z = V.dinv[i] * r;
        V.r[i] = r;
        V.z[i] = z;
        V.p[i] = z;
        v[0] = r * z;
        v[1] = r * r;
        v[2] = V.b[i] * V.b[i];
    }
    block_sum<3>(v);
    if (threadIdx.x == 0) {
        V.part[blockIdx.x] = v[0];                    // rz, parity 0
        V.part[2 * V.nb + blockIdx.x] = v[1];         // rr
        V.part[4 * V.nb + blockIdx.x] = v[2];         // bb
    }
}

// q = A p; partials of p.q
__global__ __launch_bounds__(kBlock) void k_cg_q(HeatView V) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    double v[1] = {0};
    if (i < V.N) {
        const double q = row_dot(V, i, V.p);
        V.q[i] = q;
        v[0] = V.p[i] * q;
    }
    block_sum<1>(v);
    if (threadIdx.x == 0) V.part[3 * V.nb + blockIdx.x] = v[0];
}

// alpha = rz / pq; x += alpha p; r -= alpha q; z = D^-1 r; partials of r.z (other parity), r.r
__global__ __launch_bounds__(kBlock) void k_cg_x(HeatView V, int par) {
    const double rz = sum_partials(V.part + par * V.nb, V.nb);
    const double pq = sum_partials(V.part + 3 * V.nb, V.nb);
    const double alpha = pq != 0.0 ? rz / pq : 0.0;
    const int i = blockIdx.x * kBlock + threadIdx.x;
    double v[2] = {0, 0};
    if (i < V.N) {
        V.x[i] += alpha * V.p[i];
        const double r = V.r[i] - alpha * V.q[i];
        const double z = V.dinv[i] * r;
        V.r[i] = r;
        V.z[i] = z;
        v[0] = r * z;
        v[1] = r * r;
    }
    block_sum<2>(v);
    if (threadIdx.x == 0) {
        V.part[(1 - par) * V.nb + blockIdx.x] = v[0];
        V.part[2 * V.nb + blockIdx.x] = v[1];
    }
}

// beta = rz_new / rz_old; p = z + beta p; workgroup 0 publishes r.r and b.b for the host
__global__ __launch_bounds__(kBlock) void k_cg_p(HeatView V, int par) {
    const double rz_old = sum_partials(V.part + par * V.nb, V.nb);
    const double rz_new = sum_partials(V.part + (1 - par) * V.nb, V.nb);
    const double beta = rz_old != 0.0 ? rz_new / rz_old : 0.0;
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i < V.N) V.p[i] = V.z[i] + beta * V.p[i];
    if (blockIdx.x == 0) {
        const double rr = sum_partials(V.part + 2 * V.nb, V.nb);
        const double bb = sum_partials(V.part + 4 * V.nb, V.nb);
        if (threadIdx.x == 0) {
            V.scal[0] = rr;
            V.scal[1] = bb;
        }
    }
}

// the initial r.r and b.b (before any iteration)
__global__ __launch_bounds__(kBlock) void k_cg_publish(HeatView V) {
    const double rr = sum_partials(V.part + 2 * V.nb, V.nb);
    const double bb = sum_partials(V.part + 4 * V.nb, V.nb);
    if (threadIdx.x == 0) {
        V.scal[0] = rr;
        V.scal[1] = bb;
    }
}

}  // namespace

int heat_blocks(int n) { return (n + kBlock - 1) / kBlock; }

void launch_heat_assemble(const HeatView& V, hipStream_t s) {
    const int be = heat_blocks(V.E), bn = heat_blocks(V.N), bf = heat_blocks(V.F);
    if (V.dim == 2) {
        k_heat_elem<2><<<be, kBlock, 0, s>>>(V);
        k_heat_asm<2><<<bn, kBlock, 0, s>>>(V);
        if (V.F > 0) k_heat_facet<2><<<bf, kBlock, 0, s>>>(V);
    } else {
        k_heat_elem<3><<<be, kBlock, 0, s>>>(V);
        k_heat_asm<3><<<bn, kBlock, 0, s>>>(V);
        if (V.F > 0) k_heat_facet<3><<<bf, kBlock, 0, s>>>(V);
    }
    k_heat_bc<<<bn, kBlock, 0, s>>>(V);
}

void launch_cg_init(const HeatView& V, hipStream_t s) {
    k_cg_init<<<V.nb, kBlock, 0, s>>>(V);
    k_cg_publish<<<1, kBlock, 0, s>>>(V);
}

void launch_cg_iteration(const HeatView& V, int par, hipStream_t s) {
    k_cg_q<<<V.nb, kBlock, 0, s>>>(V);
    k_cg_x<<<V.nb, kBlock, 0, s>>>(V, par);
    k_cg_p<<<V.nb, kBlock, 0, s>>>(V, par);
}

}  // namespace pfem
