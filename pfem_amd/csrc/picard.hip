// Device side of the Picard loop (PicardAlgo.cpp:31-95 with the PSPG callbacks of
// MomContEquationPSPG.inl:262-372): position update from the saved nodes, the Ax_f
// residual ||A q - b||_2 and the U / U_P relative-change residuals, batched over cases.
#include "common.hpp"
#include "nodespmv.hpp"
#include "views.hpp"

namespace pfem {

// updateNodesPositionFromSave (Mesh.cpp:1238-1277): x = x_saved + dt q_v for non-fixed nodes.
__global__ __launch_bounds__(kBlock) void k_update_pos(const PicardView* __restrict__ views,
                                                       const BlockTask* __restrict__ tasks) {
    const BlockTask task = tasks[blockIdx.x];
    if (task.case_id < 0) return;
    const PicardView& V = views[task.case_id];
    for (int i = task.block * kBlock + threadIdx.x; i < V.N; i += task.nblocks * kBlock) {
        if (V.flags[i] & NODE_FIXED) continue;
        for (int k = 0; k < V.dim; ++k) {
            const long long q = (long long)k * V.N + i;
            V.pos[q] = V.pos_save[q] + V.x[q] * V.dt;
        }
    }
}

// The copies of one Picard step selected by the task's bits (COPY_* in views.hpp); the
// step's copy kinds never touch the same array twice in one launch.
__global__ __launch_bounds__(kBlock) void k_pic_copy(const PicardView* __restrict__ views,
                                                     const BlockTask* __restrict__ tasks) {
    const BlockTask task = tasks[blockIdx.x];
    if (task.case_id < 0) return;
    const PicardView& V = views[task.case_id];
    const int op = task.pad;
    const int nf = V.dim + 1;
    for (int i = task.block * kBlock + threadIdx.x; i < V.N; i += task.nblocks * kBlock) {
        for (int k = 0; k < V.dim; ++k) {
            const long long q = (long long)k * V.N + i;
            if (op & COPY_PREP) V.pos_save[q] = V.pos[q];
            if (op & COPY_RESTORE) V.pos[q] = V.pos_save[q];
        }
        for (int k = 0; k < nf; ++k) {
            const long long q = (long long)k * V.N + i;
            if (op & COPY_PREP) {
                const double v = V.q_prev[q];
                V.x[q] = v;
                V.x_prev[q] = v;
            }
            if (op & COPY_ZERO_X) V.x[q] = 0.0;
            if (op & COPY_RESET_X) V.x[q] = V.q_prev[q];
            if (op & COPY_SAVEIT) V.x_prev[q] = V.x[q];
            if (op & COPY_COMMIT) V.q_prev[q] = V.x[q];
        }
    }
}

// ||A q - b||^2 (Ax_f cases only).
template <int L>
__global__ __launch_bounds__(kBlock) void k_res_axf(const PicardView* __restrict__ views,
                                                    const BlockTask* __restrict__ tasks) {
    constexpr int NPB = kBlock / L;
    const BlockTask task = tasks[blockIdx.x];
    if (task.case_id < 0) return;
    const PicardView& V = views[task.case_id];
    if (V.res_kind != 2) return;
    const int lane = threadIdx.x % L;
    double part[1] = {0.0};
    for (long long i = (long long)task.block * NPB + threadIdx.x / L; i < V.N; i += (long long)task.nblocks * NPB) {
        double ax[kMaxNf];
        node_spmv<L>(V.A, V.x, i, lane, ax);
        if (lane < V.A.nf) {
            const double r = pick(ax, lane) - V.b[i + (long long)lane * V.N];
            part[0] += r * r;
        }
    }
    double tot[1];
    if (reduce_last_block<1>(part, V.partials, V.counters, task.block, task.nblocks, tot) && threadIdx.x == 0)
        V.out[0] = tot[0];
}

// U / U_P cases: sum over non-free nodes of |du|^2, |u_prev|^2 (velocity) and dp^2, p_prev^2.
__global__ __launch_bounds__(kBlock) void k_res_u(const PicardView* __restrict__ views,
                                                  const BlockTask* __restrict__ tasks) {
    const BlockTask task = tasks[blockIdx.x];
    if (task.case_id < 0) return;
    const PicardView& V = views[task.case_id];
    if (V.res_kind == 2) return;
    double part[4] = {0.0, 0.0, 0.0, 0.0};
    for (int i = task.block * kBlock + threadIdx.x; i < V.N; i += task.nblocks * kBlock) {
        if (V.flags[i] & NODE_FREE) continue;
        for (int k = 0; k <= V.dim; ++k) {
            const long long q = (long long)k * V.N + i;
            const double a = V.x[q], b = V.x_prev[q];
            const int slot = (k < V.dim) ? 0 : 2;
            part[slot] += (a - b) * (a - b);
            part[slot + 1] += b * b;
        }
    }
    double tot[4];
    if (reduce_last_block<4>(part, V.partials, V.counters, task.block, task.nblocks, tot) && threadIdx.x == 0)
        for (int j = 0; j < 4; ++j) V.out[j] = tot[j];
}

void launch_pic_copy(const PicardView* v, const BlockTask* t, int nt, hipStream_t s) {
    k_pic_copy<<<nt, kBlock, 0, s>>>(v, t);
}

void launch_update_pos(const PicardView* v, const BlockTask* t, int nt, hipStream_t s) {
    k_update_pos<<<nt, kBlock, 0, s>>>(v, t);
}

void launch_res_axf(int lanes, const PicardView* v, const BlockTask* t, int nt, hipStream_t s) {
    switch (lanes) {
        case 4: k_res_axf<4><<<nt, kBlock, 0, s>>>(v, t); break;
        case 8: k_res_axf<8><<<nt, kBlock, 0, s>>>(v, t); break;
        default: k_res_axf<16><<<nt, kBlock, 0, s>>>(v, t); break;
    }
}

void launch_res_u(const PicardView* v, const BlockTask* t, int nt, hipStream_t s) { k_res_u<<<nt, kBlock, 0, s>>>(v, t); }

}  // namespace pfem
