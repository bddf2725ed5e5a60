// The fractional-step solver of the reference (Solver.id = "FracStep",
// srcs/simulation/physics/IncompNewton/MomContEquationFracStep.inl), one Picard iteration on the
// GPU as three diagonal-preconditioned conjugate-gradient solves (the reference's
// Eigen::ConjugateGradient<Lower|Upper>, MomContEquation.hpp:49, from a zero guess):
//
//   prediction   (M/dt + K) vt = F + M/dt v_prev + gammaFS D^T p_prev      (m_buildMatFracStep
//                with free rows vt = v_prev + dt g and Dirichlet rows       :7-195, m_applyBCVAppStep
//                vt = v_bc, Dirichlet columns eliminated                    :197-266)
//   pressure     L p = -(rho/dt) D vt + gammaFS L p_prev on nodes that are   (m_buildMatPcorrStep
//                neither free nor on the free surface, free rows p = 0       :268-314, :316-341)
//   correction   M dv = dt D^T (p - gammaFS p_prev), unit rows and columns  (m_buildMatVStep
//                on bound and free nodes                                     :343-429, :431-460)
//
// then v = vt + dv, positions = saved + dt v (fixed nodes stay) and the U / U_P residual over the
// nodes that are not free (:562-628).
//
// Layout: the three matrices share the node graph of the case (value k of node i's row is the
// node pair (i, nbr_idx[k])): M/dt + K as a dim x dim block per entry, M and L as scalars (M is
// applied to each velocity component: the reference's M is block diagonal with equal blocks,
// and its unit rows and columns are per node).  One gather pass per node writes all three rows
// and the prediction right-hand side from one element record per element, summing the
// contributions of the elements holding both nodes in element order (the reference's triplet
// order); the boundary conditions are applied in the same pass, row by row, in the reference's
// order (Dirichlet columns by increasing node, then component).
#include "common.hpp"
#include "fracstep.hpp"
#include "p1geom.hpp"

namespace pfem {

namespace {

constexpr int kRec = 16;  // element record: dN_i/dx_k at k (d+1) + i, K factor at 14, |e| at 15

// K factor (MomContEquation.inl:100-123): mu, or for a Bingham fluid mu + tau0 (1 - exp(-mReg g))
// / g (tau0 mReg when g < 1e-15), g = sqrt(2 sum_k L_kk^2 + sum_{k<m} (L_km + L_mk)^2) for the
// velocity gradient L of the element's node states (the current iterate).
template <int D>
__device__ double frac_k_factor(const FracView& V, const int (&nd)[D + 1], const double* gN) {
    if (V.problem != 1) return V.mu;
    double Lg[D][D];
#pragma unroll
    for (int k = 0; k < D; ++k)
#pragma unroll
        for (int m = 0; m < D; ++m) {
            double t = 0;
#pragma unroll
            for (int a = 0; a <= D; ++a) t += gN[m * (D + 1) + a] * V.v_it[(size_t)k * V.N + nd[a]];
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

template <int D>
__global__ __launch_bounds__(kBlock) void k_frac_elem(FracView V) {
    const int e = blockIdx.x * kBlock + threadIdx.x;
    if (e >= V.E) return;
    int nd[D + 1];
#pragma unroll
    for (int a = 0; a <= D; ++a) nd[a] = V.conn[(size_t)a * V.E + e];
    double x[D][D + 1];
    WGeom<D> G;
    wc_geometry<D>(V.pos, V.N, nd, x, G);
    double gN[D * (D + 1)];
#pragma unroll
    for (int k = 0; k < D; ++k)
#pragma unroll
        for (int a = 0; a <= D; ++a) gN[k * (D + 1) + a] = G.gN[k][a];
    double* rec = V.erec + (size_t)kRec * e;
#pragma unroll
    for (int t = 0; t < D * (D + 1); ++t) rec[t] = gN[t];
    rec[14] = frac_k_factor<D>(V, nd, gN);
    rec[15] = G.vol;
}

__device__ __forceinline__ bool is_bound(uint8_t f) { return (f & NODE_BOUND) != 0; }
__device__ __forceinline__ bool is_free(uint8_t f) { return (f & NODE_FREE) != 0; }

// One thread per node: its rows of M/dt + K, M and L, the prediction right-hand side, the
// boundary conditions of its rows and the three diagonal inverses.
//   M_e(a, b) = rho |e| (1 + [a = b]) / ((d+1)(d+2))                    (getM, exact quadrature)
//   K_e((a, r), (b, c)) = k |e| ([r = c] gradN_a . gradN_b + dN_a/dx_c dN_b/dx_r)   (B^T ddev B)
//   L_e(a, b) = |e| gradN_a . gradN_b,  D_e(a, (b, c)) = |e| / (d+1) dN_b/dx_c
//   F_e(a, r) = rho g_r |e| / (d+1)
template <int D>
__global__ __launch_bounds__(kBlock) void k_frac_asm(FracView V) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= V.N) return;
    const uint8_t f = V.flags[i];
    const bool bf = is_bound(f) || is_free(f);
    const double mf = V.rho / ((D + 1) * (D + 2));
    const double idt = 1.0 / V.dt;
    const int nb = V.nbr_ptr[i], deg = V.nbr_ptr[i + 1] - nb;
    double bi[D];
#pragma unroll
    for (int r = 0; r < D; ++r) bi[r] = 0.0;
    double dmk[D], dm = 1.0, dl = 1.0;
#pragma unroll
    for (int r = 0; r < D; ++r) dmk[r] = 1.0;
    for (int k = 0; k < deg; ++k) {
        const int g = nb + k;
        const int j = V.nbr_idx[g];
        double blk[D][D];
#pragma unroll
        for (int r = 0; r < D; ++r)
#pragma unroll
            for (int c = 0; c < D; ++c) blk[r][c] = 0.0;
        double m = 0.0, l = 0.0;
        for (int q = V.pei_ptr[g]; q < V.pei_ptr[g + 1]; ++q) {
            const unsigned pe = V.pei[q];
            const int e = (int)(pe >> 4), a = (int)((pe >> 2) & 3u), b = (int)(pe & 3u);
            const double* rec = V.erec + (size_t)kRec * e;
            const double vol = rec[15], kf = rec[14];
            double gg = 0;
#pragma unroll
            for (int c = 0; c < D; ++c) gg += rec[c * (D + 1) + a] * rec[c * (D + 1) + b];
            if (!bf) {
                const double me = mf * vol * (a == b ? 2.0 : 1.0);
                m += me;
                const double kv = kf * vol;
#pragma unroll
                for (int r = 0; r < D; ++r) {
                    blk[r][r] += idt * me;
#pragma unroll
                    for (int c = 0; c < D; ++c)
                        blk[r][c] += kv * ((r == c ? gg : 0.0) + rec[c * (D + 1) + a] * rec[r * (D + 1) + b]);
                }
            }
            if (!is_free(f)) l += vol * gg;
        }
        if (bf) {
#pragma unroll
            for (int r = 0; r < D; ++r)
#pragma unroll
                for (int c = 0; c < D; ++c) blk[r][c] = (j == i && r == c) ? 1.0 : 0.0;
            m = j == i ? 1.0 : 0.0;
        } else if (j != i && (is_bound(V.flags[j]) || is_free(V.flags[j]))) {
            m = 0.0;  // unit column of M (m_applyBCVStep)
        }
        if (is_free(f)) l = j == i ? 1.0 : 0.0;
        if (j == i) {
#pragma unroll
            for (int r = 0; r < D; ++r) dmk[r] = blk[r][r];
            dm = m;
            dl = l;
            if (!bf) {
                // F + M/dt v_prev + gammaFS D^T p_prev, per element in element order
                for (int q = V.pei_ptr[g]; q < V.pei_ptr[g + 1]; ++q) {
                    const unsigned pe = V.pei[q];
                    const int e = (int)(pe >> 4), a = (int)((pe >> 2) & 3u);
                    const double* rec = V.erec + (size_t)kRec * e;
                    const double vol = rec[15];
                    int nd[D + 1];
#pragma unroll
                    for (int c = 0; c <= D; ++c) nd[c] = V.conn[(size_t)c * V.E + e];
                    double dtp = 0;
#pragma unroll
                    for (int c = 0; c <= D; ++c) dtp += V.p_prev[nd[c]];
#pragma unroll
                    for (int r = 0; r < D; ++r) {
                        double mv = 0;
#pragma unroll
                        for (int c = 0; c <= D; ++c)
                            mv += idt * (mf * vol * (c == a ? 2.0 : 1.0)) * V.v_prev[(size_t)r * V.N + nd[c]];
                        const double dcoef = vol / (D + 1) * rec[r * (D + 1) + a];
                        bi[r] += V.rho * V.g[r] * vol / (D + 1);
                        bi[r] += mv;
                        bi[r] += V.gamma_fs * (dcoef * dtp);
                    }
                }
            }
        }
        double* out = V.mk + (size_t)g * D * D;
#pragma unroll
        for (int r = 0; r < D; ++r)
#pragma unroll
            for (int c = 0; c < D; ++c) out[r * D + c] = blk[r][c];
        V.ms[g] = m;
        V.lv[g] = l;
    }
    // m_applyBCVAppStep for this node's rows
    if (is_free(f) && !is_bound(f)) {
#pragma unroll
        for (int r = 0; r < D; ++r) bi[r] = V.v_prev[(size_t)r * V.N + i] + V.dt * V.g[r];
    }
    if (is_bound(f)) {
#pragma unroll
        for (int r = 0; r < D; ++r) bi[r] = (f & NODE_DIRICHLET) ? V.vbc[(size_t)r * V.N + i] : 0.0;
    } else if (!is_free(f)) {
        for (int k = 0; k < deg; ++k) {
            const int g = nb + k;
            const int j = V.nbr_idx[g];
            if (j == i || !is_bound(V.flags[j]) || !(V.flags[j] & NODE_DIRICHLET)) continue;
            double* blk = V.mk + (size_t)g * D * D;
#pragma unroll
            for (int c = 0; c < D; ++c) {
                const double vb = V.vbc[(size_t)c * V.N + j];
#pragma unroll
                for (int r = 0; r < D; ++r) {
                    bi[r] -= blk[r * D + c] * vb;
                    blk[r * D + c] = 0.0;
                }
            }
        }
    }
#pragma unroll
    for (int r = 0; r < D; ++r) {
        V.bv[(size_t)r * V.N + i] = bi[r];
        V.dinv_mk[(size_t)r * V.N + i] = dmk[r] != 0.0 ? 1.0 / dmk[r] : 1.0;
    }
#pragma unroll
    for (int r = 0; r < D; ++r) V.dinv_m[(size_t)r * V.N + i] = dm != 0.0 ? 1.0 / dm : 1.0;
    V.dinv_l[i] = dl != 0.0 ? 1.0 / dl : 1.0;
}

// One thread per node: b_p = sum_e [-(rho/dt) D_e vt + gammaFS L_e p_prev] on nodes neither free
// nor on the free surface, 0 elsewhere (the unit rows of free nodes).
template <int D>
__global__ __launch_bounds__(kBlock) void k_frac_pcorr(FracView V) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= V.N) return;
    const uint8_t f = V.flags[i];
    double bi = 0.0;
    if (!is_free(f) && !(f & NODE_ONFS)) {
        const int nb = V.nbr_ptr[i], deg = V.nbr_ptr[i + 1] - nb;
        int g = nb;
        for (int k = 0; k < deg; ++k)
            if (V.nbr_idx[nb + k] == i) g = nb + k;
        for (int q = V.pei_ptr[g]; q < V.pei_ptr[g + 1]; ++q) {
            const unsigned pe = V.pei[q];
            const int e = (int)(pe >> 4), a = (int)((pe >> 2) & 3u);
            const double* rec = V.erec + (size_t)kRec * e;
            const double vol = rec[15];
            int nd[D + 1];
#pragma unroll
            for (int c = 0; c <= D; ++c) nd[c] = V.conn[(size_t)c * V.E + e];
            double dv = 0;
#pragma unroll
            for (int r = 0; r < D; ++r)
#pragma unroll
                for (int c = 0; c <= D; ++c) dv += (vol / (D + 1) * rec[r * (D + 1) + c]) * V.vt[(size_t)r * V.N + nd[c]];
            double lp = 0;
#pragma unroll
            for (int c = 0; c <= D; ++c) {
                double gg = 0;
#pragma unroll
                for (int r = 0; r < D; ++r) gg += rec[r * (D + 1) + a] * rec[r * (D + 1) + c];
                lp += vol * gg * V.p_prev[nd[c]];
            }
            bi += -((V.rho / V.dt) * dv);
            bi += V.gamma_fs * lp;
        }
    }
    V.bp[i] = bi;
}

// One thread per node: b_v = dt sum_e D_e^T (p - gammaFS p_prev) on nodes neither free nor bound.
template <int D>
__global__ __launch_bounds__(kBlock) void k_frac_vstep(FracView V) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= V.N) return;
    const uint8_t f = V.flags[i];
    double bi[D];
#pragma unroll
    for (int r = 0; r < D; ++r) bi[r] = 0.0;
    if (!is_free(f) && !is_bound(f)) {
        const int nb = V.nbr_ptr[i], deg = V.nbr_ptr[i + 1] - nb;
        int g = nb;
        for (int k = 0; k < deg; ++k)
            if (V.nbr_idx[nb + k] == i) g = nb + k;
        for (int q = V.pei_ptr[g]; q < V.pei_ptr[g + 1]; ++q) {
            const unsigned pe = V.pei[q];
            const int e = (int)(pe >> 4), a = (int)((pe >> 2) & 3u);
            const double* rec = V.erec + (size_t)kRec * e;
            const double vol = rec[15];
            double sdp = 0;
#pragma unroll
            for (int c = 0; c <= D; ++c) {
                const int n = V.conn[(size_t)c * V.E + e];
                sdp += V.pn[n] - V.gamma_fs * V.p_prev[n];
            }
#pragma unroll
            for (int r = 0; r < D; ++r) bi[r] += V.dt * ((vol / (D + 1) * rec[r * (D + 1) + a]) * sdp);
        }
    }
#pragma unroll
    for (int r = 0; r < D; ++r) V.bv[(size_t)r * V.N + i] = bi[r];
}

// One thread per node: v = vt + dv, positions = saved + dt v for nodes not fixed, and the
// partial sums of the U / U_P residual over nodes that are not free; the previous iterate then
// takes the new one.
template <int D>
__global__ __launch_bounds__(kBlock) void k_frac_update(FracView V) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    double s[4] = {0, 0, 0, 0};
    if (i < V.N) {
        const uint8_t f = V.flags[i];
        double v[D];
#pragma unroll
        for (int r = 0; r < D; ++r) {
            v[r] = V.vt[(size_t)r * V.N + i] + V.dv[(size_t)r * V.N + i];
            V.v_it[(size_t)r * V.N + i] = v[r];
        }
        if (!(f & NODE_FIXED))
#pragma unroll
            for (int r = 0; r < D; ++r) V.pos[(size_t)r * V.N + i] = V.pos_save[(size_t)r * V.N + i] + v[r] * V.dt;
        const double p = V.pn[i];
        if (!is_free(f)) {
#pragma unroll
            for (int r = 0; r < D; ++r) {
                const double o = V.v_itp[(size_t)r * V.N + i];
                s[0] += (v[r] - o) * (v[r] - o);
                s[1] += o * o;
            }
            const double op = V.p_itp[i];
            s[2] = (p - op) * (p - op);
            s[3] = op * op;
        }
#pragma unroll
        for (int r = 0; r < D; ++r) V.v_itp[(size_t)r * V.N + i] = v[r];
        V.p_itp[i] = p;
    }
    block_sum<4>(s);
    if (threadIdx.x == 0)
#pragma unroll
        for (int k = 0; k < 4; ++k) V.res_part[(size_t)k * V.nb + blockIdx.x] = s[k];
}

// ---- conjugate gradient, diagonal preconditioner ------------------------------------------
// Reductions are fence-free: every workgroup writes its partial sums to its own slot and the
// next kernel's workgroups re-reduce them in workgroup order; r.z partials alternate between two
// arrays by iteration parity, so no kernel reads a slot that it also writes.

__device__ double sum_slots(const double* p, int nb) {
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

// row t of the system times x; a node with no graph entry (a free node: in no element) has
// unit rows in all three systems
template <int OP, int D>
__device__ __forceinline__ double cg_row(const CgView& C, int t, const double* x) {
    const int comp = OP == 0 ? 0 : t / C.N;
    const int i = t - comp * C.N;
    double s = 0;
    const int q0 = C.nbr_ptr[i], q1 = C.nbr_ptr[i + 1];
    if (q0 == q1) return x[t];
    if constexpr (OP == 2) {
        for (int q = q0; q < q1; ++q) {
            const int j = C.nbr_idx[q];
            const double* blk = C.vals + (size_t)q * D * D + comp * D;
#pragma unroll
            for (int c = 0; c < D; ++c) s += blk[c] * x[(size_t)c * C.N + j];
        }
    } else {
        const double* xc = x + (size_t)comp * C.N;
        for (int q = q0; q < q1; ++q) s += C.vals[q] * xc[C.nbr_idx[q]];
    }
    return s;
}

// r = b - A x, z = D^-1 r, p = z; partials of r.z (parity 0), r.r and b.b
template <int OP, int D>
__global__ __launch_bounds__(kBlock) void k_cg2_init(CgView C) {
    const int t = blockIdx.x * kBlock + threadIdx.x;
    double v[3] = {0, 0, 0};
    if (t < C.n) {
        const double r = C.b[t] - cg_row<OP, D>(C, t, C.x);
        const double z = C.dinv[t] * r;
        C.r[t] = r;
        C.z[t] = z;
        C.p[t] = z;
        v[0] = r * z;
        v[1] = r * r;
        v[2] = C.b[t] * C.b[t];
    }
    block_sum<3>(v);
    if (threadIdx.x == 0) {
        C.part[blockIdx.x] = v[0];
        C.part[2 * C.nb + blockIdx.x] = v[1];
        C.part[4 * C.nb + blockIdx.x] = v[2];
    }
}

// q = A p; partials of p.q
template <int OP, int D>
__global__ __launch_bounds__(kBlock) void k_cg2_q(CgView C) {
    const int t = blockIdx.x * kBlock + threadIdx.x;
    double v[1] = {0};
    if (t < C.n) {
        const double q = cg_row<OP, D>(C, t, C.p);
        C.q[t] = q;
        v[0] = C.p[t] * q;
    }
    block_sum<1>(v);
    if (threadIdx.x == 0) C.part[3 * C.nb + blockIdx.x] = v[0];
}

// alpha = rz / pq; x += alpha p; r -= alpha q; z = D^-1 r; partials of r.z (other parity), r.r
__global__ __launch_bounds__(kBlock) void k_cg2_x(CgView C, int par) {
    const double rz = sum_slots(C.part + par * C.nb, C.nb);
    const double pq = sum_slots(C.part + 3 * C.nb, C.nb);
    const double alpha = pq != 0.0 ? rz / pq : 0.0;
    const int t = blockIdx.x * kBlock + threadIdx.x;
    double v[2] = {0, 0};
    if (t < C.n) {
        C.x[t] += alpha * C.p[t];
        const double r = C.r[t] - alpha * C.q[t];
        const double z = C.dinv[t] * r;
        C.r[t] = r;
        C.z[t] = z;
        v[0] = r * z;
        v[1] = r * r;
    }
    block_sum<2>(v);
    if (threadIdx.x == 0) {
        C.part[(1 - par) * C.nb + blockIdx.x] = v[0];
        C.part[2 * C.nb + blockIdx.x] = v[1];
    }
}

// beta = rz_new / rz_old; p = z + beta p; workgroup 0 publishes r.r and b.b for the host
__global__ __launch_bounds__(kBlock) void k_cg2_p(CgView C, int par) {
    const double rz_old = sum_slots(C.part + par * C.nb, C.nb);
    const double rz_new = sum_slots(C.part + (1 - par) * C.nb, C.nb);
    const double beta = rz_old != 0.0 ? rz_new / rz_old : 0.0;
    const int t = blockIdx.x * kBlock + threadIdx.x;
    if (t < C.n) C.p[t] = C.z[t] + beta * C.p[t];
    if (blockIdx.x == 0) {
        const double rr = sum_slots(C.part + 2 * C.nb, C.nb);
        const double bb = sum_slots(C.part + 4 * C.nb, C.nb);
        if (threadIdx.x == 0) {
            C.scal[0] = rr;
            C.scal[1] = bb;
        }
    }
}

__global__ __launch_bounds__(kBlock) void k_cg2_publish(CgView C) {
    const double rr = sum_slots(C.part + 2 * C.nb, C.nb);
    const double bb = sum_slots(C.part + 4 * C.nb, C.nb);
    if (threadIdx.x == 0) {
        C.scal[0] = rr;
        C.scal[1] = bb;
    }
}

template <int OP>
void cg_init_op(const CgView& C, hipStream_t s) {
    if (C.dim == 2)
        k_cg2_init<OP, 2><<<C.nb, kBlock, 0, s>>>(C);
    else
        k_cg2_init<OP, 3><<<C.nb, kBlock, 0, s>>>(C);
}

template <int OP>
void cg_q_op(const CgView& C, hipStream_t s) {
    if (C.dim == 2)
        k_cg2_q<OP, 2><<<C.nb, kBlock, 0, s>>>(C);
    else
        k_cg2_q<OP, 3><<<C.nb, kBlock, 0, s>>>(C);
}

}  // namespace

int frac_blocks(int n) { return (n + kBlock - 1) / kBlock; }

void launch_frac_assemble(const FracView& V, hipStream_t s) {
    const int be = frac_blocks(V.E), bn = frac_blocks(V.N);
    if (V.dim == 2) {
        k_frac_elem<2><<<be, kBlock, 0, s>>>(V);
        k_frac_asm<2><<<bn, kBlock, 0, s>>>(V);
    } else {
        k_frac_elem<3><<<be, kBlock, 0, s>>>(V);
        k_frac_asm<3><<<bn, kBlock, 0, s>>>(V);
    }
}

void launch_frac_pcorr(const FracView& V, hipStream_t s) {
    const int bn = frac_blocks(V.N);
    if (V.dim == 2)
        k_frac_pcorr<2><<<bn, kBlock, 0, s>>>(V);
    else
        k_frac_pcorr<3><<<bn, kBlock, 0, s>>>(V);
}

void launch_frac_vstep(const FracView& V, hipStream_t s) {
    const int bn = frac_blocks(V.N);
    if (V.dim == 2)
        k_frac_vstep<2><<<bn, kBlock, 0, s>>>(V);
    else
        k_frac_vstep<3><<<bn, kBlock, 0, s>>>(V);
}

void launch_frac_update(const FracView& V, hipStream_t s) {
    if (V.dim == 2)
        k_frac_update<2><<<V.nb, kBlock, 0, s>>>(V);
    else
        k_frac_update<3><<<V.nb, kBlock, 0, s>>>(V);
}

void launch_cg2_init(const CgView& C, hipStream_t s) {
    if (C.op == 0)
        cg_init_op<0>(C, s);
    else if (C.op == 1)
        cg_init_op<1>(C, s);
    else
        cg_init_op<2>(C, s);
    k_cg2_publish<<<1, kBlock, 0, s>>>(C);
}

void launch_cg2_iteration(const CgView& C, int par, hipStream_t s) {
    if (C.op == 0)
        cg_q_op<0>(C, s);
    else if (C.op == 1)
        cg_q_op<1>(C, s);
    else
        cg_q_op<2>(C, s);
    k_cg2_x<<<C.nb, kBlock, 0, s>>>(C, par);
    k_cg2_p<<<C.nb, kBlock, 0, s>>>(C, par);
}

}  // namespace pfem
