// Host runtime of the PFEM gfx950 library: contexts, cases, the batched runner (device
// work tables + per-case views), the BiCGSTAB / iterative-refinement drivers, the
// Picard loop, the weakly compressible step and the C-ABI of include/pfem_gpu.h.
//
// The Picard driver restates PicardAlgo::solve (nonLinearAlgo/PicardAlgo.cpp:31-95) with
// the PSPG callbacks of MomContEquationPSPG.inl:262-372 for a *set* of cases: every
// phase (assembly, BC, Krylov, position update, residual) is one launch sequence over
// the active cases' work table, and a case leaves the set when its Picard loop ends.
// A single case is a set of one.
#include "../../include/pfem_gpu.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <memory>
#include <mutex>
#include <new>
#include <string>
#include <type_traits>
#include <vector>

#include "hostutil.hpp"
#include "launch.hpp"
#include "pattern.hpp"

using pfem::BlockTask;
using pfem::DevBuf;
using pfem::KScal;

namespace {
thread_local std::string g_err;

template <class F>
int guarded(F&& f) {
    try {
        return f();
    } catch (const pfem::Error& e) {
        g_err = e.what();
        return e.code;
    } catch (const std::bad_alloc&) {
        g_err = "host allocation failed";
        return PFEM_E_OOM;
    } catch (const std::exception& e) {
        g_err = e.what();
        return PFEM_E_INVALID;
    }
}

constexpr int kElemCap = 8192;  // workgroups per case for element / node passes
// workgroups per case in row launches (<= kMaxCaseBlocks, the partials capacity);
// PFEM_ROW_BLOCKS overrides it (tuning)
const int kRowCap = [] {
    const char* e = std::getenv("PFEM_ROW_BLOCKS");
    const int v = e ? std::atoi(e) : 0;
    return (v >= 64 && v <= pfem::kMaxCaseBlocks) ? v : 1024;
}();

int grid_for_rows(int n) { return std::max(1, std::min((n + pfem::kBlock - 1) / pfem::kBlock, kRowCap)); }
// workgroups per case in the row (Krylov) launches; KrylovView::nblk must equal it
int row_blocks(long long n) { return (int)std::max<long long>(1, std::min<long long>((n + pfem::kBlock - 1) / pfem::kBlock, kRowCap)); }

// lanes per node of the node-block SpMV: the node-graph degree (incl. the node itself) is
// ~7 for 2D Kuhn triangulations and ~15 for 3D Kuhn tetrahedralisations
int pick_lanes(double avg_deg) {
    // PFEM_SPMV_LANES (4, 8 or 16) overrides the choice (tuning)
    static const int forced = [] {
        const char* e = std::getenv("PFEM_SPMV_LANES");
        const int v = e ? std::atoi(e) : 0;
        return (v == 4 || v == 8 || v == 16) ? v : 0;
    }();
    if (forced) return forced;
    if (avg_deg > 8) return 16;
    if (avg_deg > 4) return 8;
    return 4;
}

template <class T>
pfem::NodeMat<T> node_mat(const pfem::Pattern& P, int dim, const T* vals) {
    pfem::NodeMat<T> A{};
    A.N = P.N;
    A.nf = dim + 1;
    A.mom_nnz = P.mom_nnz;
    A.info = P.node_info.as<int4>();
    A.nbr_ptr = P.nbr_ptr.as<int>();
    A.nbr_idx = P.nbr_idx.as<int>();
    A.vals = vals;
    return A;
}
}  // namespace

namespace pfem {
struct Runner;
}

struct pfem_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::unique_ptr<pfem::Runner> runner;
    std::mutex mu;
};

struct pfem_case {
    pfem_ctx* ctx = nullptr;
    pfem_case_desc d{};
    hipStream_t stream = nullptr;
    int dim = 0, N = 0, E = 0, n = 0;
    bool has_mesh = false;
    pfem::Pattern pat;
    DevBuf pos, pos_save, dir_vals;
    // PSPG system, Picard state and Krylov workspace
    DevBuf q_prev, x, q_itprev, vals, b, vals32, binv, binv32, kv64, kv32;
    DevBuf partials, counters, scal, out;
    // WC state and accumulators
    DevBuf wstate, wacc, wdt2;
    pfem::PspgView pv{};
    pfem::PicardView picv{};
    pfem::KrylovView<double> k64{};
    pfem::KrylovView<float> k32{};
    pfem::RefineView rv{};
    pfem::WcView<double> wc64{};
    pfem::WcView<float> wc32{};
    KScal* h_scal = nullptr;  // pinned [2]: FP64 solve, FP32 inner solve
    double* h_out = nullptr;  // pinned [4]
    int lanes = 8;
    std::unique_ptr<pfem::Runner> runner;
    ~pfem_case();
};

namespace pfem {

// Device arrays of per-case views plus the work tables of one set of cases.
struct Runner {
    hipStream_t s = nullptr;
    std::vector<pfem_case*> act;
    DevBuf v_pspg, v_pic, v_k64, v_k32, v_ref, v_wc64, v_wc32;
    DevBuf t_elem[2], t_asm[2], t_node, t_row;
    int nt_elem[2] = {0, 0}, nt_asm[2] = {0, 0}, nt_node = 0, nt_row = 0;
    int lanes = 8;
    int par = 0;  // slot parity of the next BiCGSTAB iteration launch
    hipEvent_t ev[6] = {};
    // live timing of the first SpMV kernel of each BiCGSTAB iteration (single-case runs)
    bool prof = false;
    hipEvent_t pev[128] = {};
    double prof_ms[2] = {0, 0};
    long long prof_count[2] = {0, 0};
    explicit Runner(hipStream_t st) : s(st) {
        for (auto& e : ev) PFEM_HIP(hipEventCreate(&e));
    }
    ~Runner() {
        for (auto& e : ev)
            if (e) (void)hipEventDestroy(e);
        for (auto& e : pev)
            if (e) (void)hipEventDestroy(e);
    }
    void enable_profiling(bool on) {
        if (on && !pev[0])  // timing-only events: no system-scope fence (no L2 write-back) on record
            for (auto& e : pev) PFEM_HIP(hipEventCreateWithFlags(&e, hipEventDisableSystemFence));
        prof = on;
        if (on) prof_ms[0] = prof_ms[1] = 0, prof_count[0] = prof_count[1] = 0;
    }

    template <class V>
    void upload_views(DevBuf& dst, V pfem_case::*member) {
        std::vector<V> h(act.size());
        for (size_t i = 0; i < act.size(); ++i) h[i] = act[i]->*member;
        dst.reserve(sizeof(V) * std::max<size_t>(h.size(), 1));
        PFEM_HIP(hipMemcpyAsync(dst.p, h.data(), sizeof(V) * h.size(), hipMemcpyHostToDevice, s));
        PFEM_HIP(hipStreamSynchronize(s));
    }

    static int upload_tasks(DevBuf& dst, const std::vector<BlockTask>& t, hipStream_t s) {
        dst.reserve(sizeof(BlockTask) * std::max<size_t>(t.size(), 1));
        if (!t.empty()) PFEM_HIP(hipMemcpyAsync(dst.p, t.data(), sizeof(BlockTask) * t.size(), hipMemcpyHostToDevice, s));
        PFEM_HIP(hipStreamSynchronize(s));
        return (int)t.size();
    }

    static void append(std::vector<BlockTask>& t, int cid, long long work, int cap) {
        long long nb = (work + kBlock - 1) / kBlock;
        nb = std::max<long long>(1, std::min<long long>(nb, cap));
        for (int b = 0; b < nb; ++b) t.push_back(BlockTask{cid, b, (int)nb, 0});
    }

    // Makes `cs` the active set: uploads every view type and builds the work tables.
    void set_active(const std::vector<pfem_case*>& cs) {
        act = cs;
        upload_views(v_pspg, &pfem_case::pv);
        upload_views(v_pic, &pfem_case::picv);
        upload_views(v_k64, &pfem_case::k64);
        upload_views(v_k32, &pfem_case::k32);
        upload_views(v_ref, &pfem_case::rv);
        upload_views(v_wc64, &pfem_case::wc64);
        upload_views(v_wc32, &pfem_case::wc32);
        std::vector<BlockTask> e2, e3, a2, a3, nodes, rows;
        double gnnz = 0, gnodes = 0;
        for (size_t i = 0; i < cs.size(); ++i) {
            pfem_case* c = cs[i];
            append(c->dim == 2 ? e2 : e3, (int)i, c->E, kElemCap);
            // node-major assembly: kAsmLanes (16) lanes per node
            append(c->dim == 2 ? a2 : a3, (int)i, (long long)c->N * 16, kElemCap);
            append(nodes, (int)i, c->N, kElemCap);
            for (int b = 0, nb = row_blocks(c->n); b < nb; ++b) rows.push_back(BlockTask{(int)i, b, nb, 0});
            if (c->d.kind == PFEM_KIND_PSPG) {
                gnnz += (double)c->pat.graph_nnz;
                gnodes += c->N;
            }
        }
        nt_elem[0] = upload_tasks(t_elem[0], e2, s);
        nt_elem[1] = upload_tasks(t_elem[1], e3, s);
        nt_asm[0] = upload_tasks(t_asm[0], a2, s);
        nt_asm[1] = upload_tasks(t_asm[1], a3, s);
        nt_node = upload_tasks(t_node, nodes, s);
        nt_row = upload_tasks(t_row, rows, s);
        lanes = pick_lanes(gnodes > 0 ? gnnz / gnodes : 15.0);
    }

    // Refresh the views only (same set, changed per-step scalars such as dt).
    void refresh_views() {
        upload_views(v_pspg, &pfem_case::pv);
        upload_views(v_pic, &pfem_case::picv);
        upload_views(v_wc64, &pfem_case::wc64);
        upload_views(v_wc32, &pfem_case::wc32);
    }

    const BlockTask* te(int k) const { return t_elem[k].as<BlockTask>(); }
};

// ---------------------------------------------------------------------------------------
// PSPG phases over the active set
// ---------------------------------------------------------------------------------------
static void assemble_active(Runner& R) {
    const PspgView* v = R.v_pspg.as<PspgView>();
    if (R.nt_asm[0]) launch_pspg_assemble(2, v, R.t_asm[0].as<BlockTask>(), R.nt_asm[0], R.s);
    if (R.nt_asm[1]) launch_pspg_assemble(3, v, R.t_asm[1].as<BlockTask>(), R.nt_asm[1], R.s);
    launch_pspg_bc(v, R.t_node.as<BlockTask>(), R.nt_node, R.s);
    launch_pspg_precond(v, R.t_row.as<BlockTask>(), R.nt_row, R.s);
    PFEM_HIP(hipGetLastError());
}

static float ev_ms(hipEvent_t a, hipEvent_t b) {
    float ms = 0;
    if (hipEventElapsedTime(&ms, a, b) != hipSuccess) ms = 0;
    return ms;
}

struct KryResult {
    int status = PFEM_OK;
    int iters = 0;
    int restarts = 0;
    double rr = 0, bb = 0;
};

// BiCGSTAB over the active set with device-side scalars; the host polls the done flags.
// verify: a converged case is re-initialised once so that convergence is confirmed on
// the true residual b - A x.
template <class T>
static void krylov_run(Runner& R, const std::vector<pfem_case*>& cs, bool verify, std::vector<KryResult>& res) {
    constexpr int which = std::is_same<T, double>::value ? 0 : 1;
    const size_t nc = cs.size();
    res.assign(nc, KryResult{});
    if (nc == 0) return;
    if (R.act != cs) R.set_active(cs);
    for (pfem_case* c : cs) {
        KScal& k = c->h_scal[which];
        std::memset(&k, 0, sizeof(KScal));
        k.rtol = which == 0 ? c->d.krylov_rtol : c->d.inner_rtol;
        k.maxit = which == 0 ? c->d.krylov_max_iter : std::max(1, c->d.inner_max_iter);
        k.req_init = 1;
        PFEM_HIP(hipMemcpyAsync(c->scal.as<KScal>() + which, &k, sizeof(KScal), hipMemcpyHostToDevice, R.s));
    }
    const KrylovView<T>* v = which == 0 ? (const KrylovView<T>*)R.v_k64.p : (const KrylovView<T>*)R.v_k32.p;
    const BlockTask* t = R.t_row.as<BlockTask>();
    // init_fin writes the slot the next iteration reads as its previous one
    launch_kry_init<T>(R.lanes, v, t, R.nt_row, (int)nc, R.par ^ 1, R.s);
    std::vector<char> fin(nc, 0), verifying(nc, 0);
    int chunk = 4;
    const bool prof = R.prof && nc == 1;
    int prof_chunk = 0, prof_before = 0;
    for (;;) {
        for (pfem_case* c : cs)
            PFEM_HIP(hipMemcpyAsync(&c->h_scal[which], c->scal.as<KScal>() + which, sizeof(KScal),
                                    hipMemcpyDeviceToHost, R.s));
        PFEM_HIP(hipStreamSynchronize(R.s));
        if (prof_chunk > 0) {
            // only iterations that did work (the case may converge inside the chunk)
            const int worked = std::min(prof_chunk, cs[0]->h_scal[which].iters - prof_before);
            for (int it = 0; it < worked; ++it) {
                R.prof_ms[which] += ev_ms(R.pev[2 * it], R.pev[2 * it + 1]);
                R.prof_count[which]++;
            }
            prof_chunk = 0;
        }
        bool need_init = false, all_final = true;
        for (size_t i = 0; i < nc; ++i) {
            if (fin[i]) continue;
            pfem_case* c = cs[i];
            KScal& k = c->h_scal[which];
            res[i].iters = k.iters;
            res[i].rr = k.rr;
            res[i].bb = k.bb;
            bool again = false;
            if (k.done == 0) {
                all_final = false;
                continue;
            } else if (k.done == 1) {
                if (!verify || verifying[i]) {
                    fin[i] = 1;
                    res[i].status = PFEM_OK;
                } else {
                    verifying[i] = 1;
                    again = true;
                }
            } else if (k.done == 2) {
                if (res[i].restarts < 8 && k.iters < k.maxit) {
                    res[i].restarts++;
                    verifying[i] = 0;
                    again = true;
                } else {
                    fin[i] = 1;
                    res[i].status = PFEM_NAN;
                }
            } else {
                fin[i] = 1;
                res[i].status = (k.bb > 0 && std::isfinite(k.rr)) ? PFEM_NOT_CONVERGED : PFEM_NAN;
            }
            if (again) {
                k.done = 0;
                k.req_init = 1;
                PFEM_HIP(hipMemcpyAsync(c->scal.as<KScal>() + which, &k, sizeof(KScal), hipMemcpyHostToDevice, R.s));
                need_init = true;
                all_final = false;
            }
        }
        if (all_final) break;
        if (need_init) {
            launch_kry_init<T>(R.lanes, v, t, R.nt_row, (int)nc, R.par ^ 1, R.s);
            continue;
        }
        if (prof) {
            prof_chunk = chunk;
            prof_before = cs[0]->h_scal[which].iters;
        }
        launch_kry_iters<T>(R.lanes, chunk, v, t, R.nt_row, R.par, R.s, prof ? R.pev : nullptr);
        R.par = (R.par + chunk) & 1;
        PFEM_HIP(hipGetLastError());
        chunk = std::min(chunk * 2, 64);
    }
}

// Linear solve of every active case (FP64 BiCGSTAB, or FP32 inner BiCGSTAB inside FP64
// iterative refinement).  Solution in c->x (initial guess on entry).
static void solve_active(Runner& R, const std::vector<pfem_case*>& cs, std::vector<pfem_solve_stats>& st) {
    const size_t nc = cs.size();
    st.assign(nc, pfem_solve_stats{});
    std::vector<pfem_case*> g64, gmx;
    std::vector<size_t> i64, imx;
    for (size_t i = 0; i < nc; ++i) {
        if (cs[i]->d.precision == PFEM_PREC_MIXED) {
            gmx.push_back(cs[i]);
            imx.push_back(i);
        } else {
            g64.push_back(cs[i]);
            i64.push_back(i);
        }
    }
    if (!g64.empty()) {
        std::vector<KryResult> kr;
        krylov_run<double>(R, g64, true, kr);
        for (size_t j = 0; j < g64.size(); ++j) {
            auto& s = st[i64[j]];
            s.status = kr[j].status;
            s.krylov_iters = kr[j].iters;
            s.restarts = kr[j].restarts;
            s.relres = kr[j].bb > 0 ? std::sqrt(kr[j].rr / kr[j].bb) : 0.0;
        }
    }
    if (!gmx.empty()) {
        // FP64 iterative refinement around capped FP32 inner BiCGSTAB solves.  A case whose
        // true FP64 residual stops decreasing (the FP32 recurrence can lose accuracy on
        // badly conditioned systems) finishes with FP64 BiCGSTAB from its current iterate.
        std::vector<pfem_case*> act = gmx;
        std::vector<size_t> idx(gmx.size());
        for (size_t j = 0; j < gmx.size(); ++j) idx[j] = imx[j];
        std::vector<double> best(nc, std::numeric_limits<double>::infinity());
        std::vector<int> stall(nc, 0);
        std::vector<pfem_case*> fallback;
        std::vector<size_t> fidx;
        const int max_ref = 40;
        for (int ref = 0; !act.empty(); ++ref) {
            if (R.act != act) R.set_active(act);
            launch_ir_residual(R.lanes, R.v_ref.as<RefineView>(), R.t_row.as<BlockTask>(), R.nt_row, R.s);
            for (pfem_case* c : act)
                PFEM_HIP(hipMemcpyAsync(c->h_out, c->out.p, 2 * sizeof(double), hipMemcpyDeviceToHost, R.s));
            PFEM_HIP(hipStreamSynchronize(R.s));
            std::vector<pfem_case*> next;
            std::vector<size_t> nidx;
            for (size_t j = 0; j < act.size(); ++j) {
                pfem_case* c = act[j];
                const size_t i = idx[j];
                auto& s = st[i];
                const double rr = c->h_out[0], bb = c->h_out[1];
                const double relres = bb > 0 ? std::sqrt(rr / bb) : 0.0;
                s.relres = relres;
                s.refinements = ref;
                if (relres <= c->d.krylov_rtol) {
                    s.status = PFEM_OK;
                    continue;
                }
                if (std::isfinite(relres) && relres < 0.5 * best[i]) stall[i] = 0;
                else stall[i]++;
                if (std::isfinite(relres)) best[i] = std::min(best[i], relres);
                if (stall[i] >= 3 || !std::isfinite(relres) || ref >= max_ref ||
                    s.krylov_iters >= c->d.krylov_max_iter) {
                    fallback.push_back(c);
                    fidx.push_back(i);
                } else {
                    next.push_back(c);
                    nidx.push_back(i);
                }
            }
            if (next.empty()) break;
            std::vector<KryResult> kr;
            krylov_run<float>(R, next, false, kr);
            for (size_t j = 0; j < next.size(); ++j) {
                st[nidx[j]].krylov_iters += kr[j].iters;
                st[nidx[j]].restarts += kr[j].restarts;
            }
            if (R.act != next) R.set_active(next);
            launch_ir_update(R.v_ref.as<RefineView>(), R.t_row.as<BlockTask>(), R.nt_row, R.s);
            PFEM_HIP(hipGetLastError());
            act = next;
            idx = nidx;
        }
        if (!fallback.empty()) {
            // a non-finite iterate restarts from zero
            for (size_t j = 0; j < fallback.size(); ++j)
                if (!std::isfinite(st[fidx[j]].relres))
                    PFEM_HIP(hipMemsetAsync(fallback[j]->x.p, 0, sizeof(double) * (size_t)fallback[j]->n, R.s));
            std::vector<KryResult> kr;
            krylov_run<double>(R, fallback, true, kr);
            for (size_t j = 0; j < fallback.size(); ++j) {
                auto& s = st[fidx[j]];
                s.status = kr[j].status;
                s.krylov_iters += kr[j].iters;
                s.restarts += kr[j].restarts + 1;
                s.relres = kr[j].bb > 0 ? std::sqrt(kr[j].rr / kr[j].bb) : 0.0;
            }
        }
    }
}

// One Picard time step for each case of `cs` (PicardAlgo.cpp:31-95).
static void pspg_step_set(Runner& R, const std::vector<pfem_case*>& cs, const std::vector<double>& dts,
                          std::vector<pfem_step_stats>& out) {
    const size_t nc = cs.size();
    out.assign(nc, pfem_step_stats{});
    for (size_t i = 0; i < nc; ++i) {
        pfem_case* c = cs[i];
        PFEM_REQUIRE(c->has_mesh && c->d.kind == PFEM_KIND_PSPG, PFEM_E_STATE, "case has no PSPG mesh");
        c->pv.dt = dts[i];
        c->picv.dt = dts[i];
        const size_t nb = sizeof(double) * (size_t)c->n;
        // prepare (MomContEquationPSPG.inl:264-277): save nodes; node states = qPrev
        PFEM_HIP(hipMemcpyAsync(c->pos_save.p, c->pos.p, sizeof(double) * (size_t)c->dim * c->N,
                                hipMemcpyDeviceToDevice, R.s));
        PFEM_HIP(hipMemcpyAsync(c->x.p, c->q_prev.p, nb, hipMemcpyDeviceToDevice, R.s));
        PFEM_HIP(hipMemcpyAsync(c->q_itprev.p, c->q_prev.p, nb, hipMemcpyDeviceToDevice, R.s));
    }
    R.set_active(cs);
    PFEM_HIP(hipEventRecord(R.ev[0], R.s));
    assemble_active(R);
    PFEM_HIP(hipEventRecord(R.ev[1], R.s));
    for (auto& o : out) o.assemblies = 1;
    std::vector<int> iter(nc, 0);
    std::vector<double> res(nc, std::numeric_limits<double>::max());
    std::vector<char> live(nc, 1);
    bool first = true;
    auto restore = [&](size_t i) {
        pfem_case* c = cs[i];
        PFEM_HIP(hipMemcpyAsync(c->pos.p, c->pos_save.p, sizeof(double) * (size_t)c->dim * c->N,
                                hipMemcpyDeviceToDevice, R.s));
    };
    for (;;) {
        std::vector<pfem_case*> act;
        std::vector<size_t> idx;
        for (size_t i = 0; i < nc; ++i) {
            if (!live[i]) continue;
            if (!(res[i] > cs[i]->d.min_res)) {  // while (res > minRes)
                live[i] = 0;
                out[i].status = PFEM_OK;
                PFEM_HIP(hipMemcpyAsync(cs[i]->q_prev.p, cs[i]->x.p, sizeof(double) * (size_t)cs[i]->n,
                                        hipMemcpyDeviceToDevice, R.s));
                continue;
            }
            if (iter[i] > cs[i]->d.max_iter) {  // iterCount > maxIter
                live[i] = 0;
                out[i].status = PFEM_NOT_CONVERGED;
                restore(i);
                continue;
            }
            act.push_back(cs[i]);
            idx.push_back(i);
        }
        if (act.empty()) break;
        // solve callback (MomContEquationPSPG.inl:278-313)
        if (R.act != act) R.set_active(act);
        PFEM_HIP(hipEventRecord(R.ev[2], R.s));
        std::vector<pfem_solve_stats> ss;
        solve_active(R, act, ss);
        PFEM_HIP(hipEventRecord(R.ev[3], R.s));
        std::vector<pfem_case*> ok;
        std::vector<size_t> oidx;
        for (size_t j = 0; j < act.size(); ++j) {
            const size_t i = idx[j];
            out[i].krylov_iters += ss[j].krylov_iters;
            out[i].max_relres = std::max(out[i].max_relres, ss[j].relres);
            if (ss[j].status != PFEM_OK && ss[j].status != PFEM_NOT_CONVERGED) {
                live[i] = 0;
                out[i].status = ss[j].status;
                restore(i);
            } else if (ss[j].status == PFEM_NOT_CONVERGED) {
                live[i] = 0;
                out[i].status = PFEM_NOT_CONVERGED;
                restore(i);
            } else {
                ok.push_back(act[j]);
                oidx.push_back(i);
            }
        }
        if (ok.empty()) continue;
        if (R.act != ok) R.set_active(ok);
        // states = q; x = x_saved + dt q (updateNodesPositionFromSave); re-assemble
        launch_update_pos(R.v_pic.as<PicardView>(), R.t_node.as<BlockTask>(), R.nt_node, R.s);
        PFEM_HIP(hipEventRecord(R.ev[4], R.s));
        assemble_active(R);
        for (size_t i : oidx) out[i].assemblies++;
        // residual callback (MomContEquationPSPG.inl:314-372)
        launch_res_axf(R.lanes, R.v_pic.as<PicardView>(), R.t_row.as<BlockTask>(), R.nt_row, R.s);
        // both residual kernels reduce through the per-case partials (kMaxCaseBlocks slots),
        // so they run over the row table, whose per-case block count is capped at kRowCap
        launch_res_u(R.v_pic.as<PicardView>(), R.t_row.as<BlockTask>(), R.nt_row, R.s);
        PFEM_HIP(hipGetLastError());
        for (pfem_case* c : ok) {
            PFEM_HIP(hipMemcpyAsync(c->h_out, c->out.p, 4 * sizeof(double), hipMemcpyDeviceToHost, R.s));
            PFEM_HIP(hipMemcpyAsync(c->q_itprev.p, c->x.p, sizeof(double) * (size_t)c->n, hipMemcpyDeviceToDevice,
                                    R.s));
        }
        PFEM_HIP(hipEventRecord(R.ev[5], R.s));
        PFEM_HIP(hipStreamSynchronize(R.s));
        if (nc == 1) {
            auto& o = out[0];
            if (first) o.t_assemble += 1e-3 * ev_ms(R.ev[0], R.ev[1]);
            o.t_solve += 1e-3 * ev_ms(R.ev[2], R.ev[3]);
            o.t_update += 1e-3 * ev_ms(R.ev[3], R.ev[4]);
            o.t_residual += 1e-3 * ev_ms(R.ev[4], R.ev[5]);
        }
        first = false;
        for (size_t j = 0; j < ok.size(); ++j) {
            const size_t i = oidx[j];
            pfem_case* c = ok[j];
            const double* o = c->h_out;
            double r;
            if (c->d.residual == PFEM_RES_AX_F) {
                r = std::sqrt(o[0]);
            } else {
                const double big = std::numeric_limits<double>::max();
                const double rv = (o[1] == 0) ? big : std::sqrt(o[0] / o[1]);
                const double rp = (c->d.residual == PFEM_RES_U_P) ? ((o[3] == 0) ? big : std::sqrt(o[2] / o[3])) : 0.0;
                r = std::max(rv, rp);
            }
            res[i] = r;
            out[i].residual = r;
            if (std::isnan(r)) {
                live[i] = 0;
                out[i].status = PFEM_NAN;
                restore(i);
                out[i].picard_iters = iter[i] + 1;
                continue;
            }
            iter[i]++;
            out[i].picard_iters = iter[i];
        }
    }
    PFEM_HIP(hipStreamSynchronize(R.s));
}

// One WC step for each case (grouped by state precision).
template <class S>
static void wc_step_group(Runner& R, const std::vector<pfem_case*>& cs) {
    if (cs.empty()) return;
    R.set_active(cs);
    const WcView<S>* v = std::is_same<S, double>::value ? (const WcView<S>*)R.v_wc64.p : (const WcView<S>*)R.v_wc32.p;
    launch_wc_step<S>(v, R.t_node.as<BlockTask>(), R.nt_node, R.te(0), R.nt_elem[0], R.te(1), R.nt_elem[1], R.s);
    PFEM_HIP(hipGetLastError());
}

}  // namespace pfem

// =========================================================================================
// case construction
// =========================================================================================
pfem_case::~pfem_case() {
    runner.reset();
    if (h_scal) (void)hipHostFree(h_scal);
    if (h_out) (void)hipHostFree(h_out);
    if (stream) (void)hipStreamDestroy(stream);
}

namespace {

void fill_views(pfem_case* c) {
    using namespace pfem;
    const int N = c->N, dim = c->dim;
    const long long n = c->n;
    auto& P = c->pat;
    // PSPG
    PspgView& pv = c->pv;
    pv = PspgView{};
    pv.dim = dim;
    pv.N = N;
    pv.E = c->E;
    pv.mom_nnz = P.mom_nnz;
    pv.conn = P.conn.as<int>();
    pv.pos = c->pos.as<double>();
    pv.q_prev = c->q_prev.as<double>();
    pv.q_iter = c->x.as<double>();
    pv.node_info = P.node_info.as<int4>();
    pv.nbr_ptr = P.nbr_ptr.as<int>();
    pv.nbr_idx = P.nbr_idx.as<int>();
    pv.pei_ptr = P.pei_ptr.as<int>();
    pv.pei = P.pei.as<uint32_t>();
    pv.dir_vals = c->dir_vals.as<double>();
    pv.vals = c->vals.as<double>();
    pv.b = c->b.as<double>();
    const bool mixed = c->d.precision == PFEM_PREC_MIXED;
    pv.vals32 = mixed ? c->vals32.as<float>() : nullptr;
    pv.diag_pos = P.diag_pos.as<int>();
    pv.binv = c->binv.as<double>();
    pv.binv32 = mixed ? c->binv32.as<float>() : nullptr;
    pv.nnz = P.nnz;
    pv.rho = c->d.rho;
    pv.mu = c->d.mu;
    pv.dt = 0;
    for (int k = 0; k < 3; ++k) pv.g[k] = c->d.body_force[k];
    // Picard
    PicardView& pc = c->picv;
    pc = PicardView{};
    pc.dim = dim;
    pc.N = N;
    pc.n = (int)n;
    pc.res_kind = c->d.residual;
    pc.flags = P.flags.as<uint8_t>();
    pc.pos = c->pos.as<double>();
    pc.pos_save = c->pos_save.as<double>();
    pc.x = c->x.as<double>();
    pc.x_prev = c->q_itprev.as<double>();
    pc.A = node_mat<double>(P, dim, c->vals.as<double>());
    pc.b = c->b.as<double>();
    pc.partials = c->partials.as<double>();
    pc.counters = c->counters.as<unsigned>();
    pc.out = c->out.as<double>();
    // Krylov FP64: x is the Picard iterate
    auto& k = c->k64;
    k = KrylovView<double>{};
    k.n = (int)n;
    k.x_zero = 0;
    k.nodes = N;
    k.nf = dim + 1;
    k.nblk = row_blocks(n);
    k.A = node_mat<double>(P, dim, c->vals.as<double>());
    k.binv = c->binv.as<double>();
    k.b = c->b.as<double>();
    k.x = c->x.as<double>();
    double* w = c->kv64.as<double>();
    k.r = w;
    k.rh = w + n;
    k.p = w + 2 * n;
    k.v = w + 3 * n;
    k.s = w + 4 * n;
    k.t = w + 5 * n;
    k.ph = w + 6 * n;
    k.sh = w + 7 * n;
    k.partials = c->partials.as<double>();
    k.counters = c->counters.as<unsigned>();
    k.sc = c->scal.as<KScal>();
    // Krylov FP32 inner solve: b32 = r32 (refinement residual), x32 = correction
    auto& f = c->k32;
    f = KrylovView<float>{};
    if (mixed) {
        float* u = c->kv32.as<float>();
        f.n = (int)n;
        f.x_zero = 1;
        f.nodes = N;
        f.nf = dim + 1;
        f.nblk = row_blocks(n);
        f.A = node_mat<float>(P, dim, c->vals32.as<float>());
        f.binv = c->binv32.as<float>();
        f.b = u;
        f.x = u + n;
        f.r = u + 2 * n;
        f.rh = u + 3 * n;
        f.p = u + 4 * n;
        f.v = u + 5 * n;
        f.s = u + 6 * n;
        f.t = u + 7 * n;
        f.ph = u + 8 * n;
        f.sh = u + 9 * n;
        f.partials = c->partials.as<double>();
        f.counters = c->counters.as<unsigned>();
        f.sc = c->scal.as<KScal>() + 1;
        auto& r = c->rv;
        r = RefineView{};
        r.n = (int)n;
        r.A = node_mat<double>(P, dim, c->vals.as<double>());
        r.b = c->b.as<double>();
        r.x = c->x.as<double>();
        r.r = c->kv64.as<double>();
        r.r32 = u;
        r.d32 = u + n;
        r.partials = c->partials.as<double>();
        r.counters = c->counters.as<unsigned>();
        r.out = c->out.as<double>();
    }
}

void fill_wc_views(pfem_case* c) {
    using namespace pfem;
    const int N = c->N, dim = c->dim;
    auto set = [&](auto& V, auto* base) {
        using S = std::remove_pointer_t<decltype(base)>;
        V = WcView<S>{};
        V.dim = dim;
        V.N = N;
        V.E = c->E;
        V.conn = c->pat.conn.as<int>();
        V.flags = c->pat.flags.as<uint8_t>();
        V.pos = c->pos.as<double>();
        // layout of pfem_case_set_state: v (dim x N), p (N), rho (N), a (dim x N)
        V.v = base;
        V.p = base + (size_t)dim * N;
        V.rho = base + (size_t)(dim + 1) * N;
        V.a = base + (size_t)(dim + 2) * N;
        V.dir_vals = c->dir_vals.as<double>();
        double* acc = c->wacc.as<double>();
        V.m_cont = acc;
        V.f_cont = acc + N;
        V.m_mom = acc + 2 * (size_t)N;
        V.f_mom = acc + 3 * (size_t)N;
        V.dt2 = c->wdt2.as<unsigned long long>();
        V.mu = c->d.mu;
        V.K0 = c->d.K0;
        V.K0p = c->d.K0p;
        V.rho_star = c->d.rho_star;
        for (int k = 0; k < 3; ++k) V.g[k] = c->d.body_force[k];
        V.meduri = c->d.meduri;
    };
    c->wc64 = WcView<double>{};
    c->wc32 = WcView<float>{};
    if (c->d.wc_state == PFEM_STATE_FP32) set(c->wc32, c->wstate.as<float>());
    else set(c->wc64, c->wstate.as<double>());
}

void require_case(pfem_case* c) { PFEM_REQUIRE(c != nullptr, PFEM_E_INVALID, "null case handle"); }
void require_mesh(pfem_case* c) {
    require_case(c);
    PFEM_REQUIRE(c->has_mesh, PFEM_E_STATE, "no mesh set on this case");
}

pfem::Runner& case_runner(pfem_case* c) {
    if (!c->runner) c->runner = std::make_unique<pfem::Runner>(c->stream);
    return *c->runner;
}

size_t wc_state_size(pfem_case* c) { return c->d.wc_state == PFEM_STATE_FP32 ? sizeof(float) : sizeof(double); }

}  // namespace

// =========================================================================================
// C-ABI
// =========================================================================================
extern "C" {

const char* pfem_last_error(void) { return g_err.c_str(); }
int pfem_version(void) { return 1; }

int pfem_ctx_create(int32_t device, pfem_ctx** out) {
    return guarded([&] {
        PFEM_REQUIRE(out != nullptr, PFEM_E_INVALID, "null output pointer");
        int count = 0;
        PFEM_HIP(hipGetDeviceCount(&count));
        PFEM_REQUIRE(device >= 0 && device < count, PFEM_E_INVALID, "device index out of range");
        PFEM_HIP(hipSetDevice(device));
        auto ctx = std::make_unique<pfem_ctx>();
        ctx->device = device;
        PFEM_HIP(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking));
        ctx->runner = std::make_unique<pfem::Runner>(ctx->stream);
        *out = ctx.release();
        return PFEM_OK;
    });
}

int pfem_ctx_destroy(pfem_ctx* ctx) {
    return guarded([&] {
        if (!ctx) return PFEM_OK;
        (void)hipSetDevice(ctx->device);
        (void)hipStreamSynchronize(ctx->stream);
        ctx->runner.reset();
        (void)hipStreamDestroy(ctx->stream);
        delete ctx;
        return PFEM_OK;
    });
}

int pfem_case_create(pfem_ctx* ctx, const pfem_case_desc* desc, pfem_case** out) {
    return guarded([&] {
        PFEM_REQUIRE(ctx && desc && out, PFEM_E_INVALID, "null argument");
        const auto& d = *desc;
        PFEM_REQUIRE(d.dim == 2 || d.dim == 3, PFEM_E_INVALID, "dim must be 2 or 3");
        PFEM_REQUIRE(d.kind == PFEM_KIND_PSPG || d.kind == PFEM_KIND_WC, PFEM_E_INVALID, "unknown case kind");
        PFEM_REQUIRE(std::fabs(d.gamma) < 1e-15, PFEM_E_UNSUPPORTED,
                     "surface tension (gamma > 0) is not supported: the FST facet term is not implemented");
        PFEM_REQUIRE(d.mu >= 0, PFEM_E_INVALID, "mu must be >= 0");
        if (d.kind == PFEM_KIND_PSPG) {
            PFEM_REQUIRE(d.rho > 0, PFEM_E_INVALID, "rho must be > 0");
            PFEM_REQUIRE(d.residual >= 0 && d.residual <= 2, PFEM_E_INVALID, "unknown residual kind");
            PFEM_REQUIRE(d.max_iter >= 0, PFEM_E_INVALID, "maxIter must be >= 0");
            PFEM_REQUIRE(d.precision == PFEM_PREC_FP64 || d.precision == PFEM_PREC_MIXED, PFEM_E_INVALID,
                         "unknown precision");
            PFEM_REQUIRE(d.krylov_rtol > 0 && d.krylov_max_iter > 0, PFEM_E_INVALID, "invalid Krylov settings");
            PFEM_REQUIRE(d.precision != PFEM_PREC_MIXED || (d.inner_rtol > 0 && d.inner_max_iter > 0), PFEM_E_INVALID,
                         "inner_rtol and inner_max_iter must be > 0");
        } else {
            PFEM_REQUIRE(d.K0 > 0 && d.K0p > 0 && d.rho_star > 0, PFEM_E_INVALID, "invalid Tait parameters");
            PFEM_REQUIRE(d.wc_state == PFEM_STATE_FP64 || d.wc_state == PFEM_STATE_FP32, PFEM_E_INVALID,
                         "unknown WC state precision");
        }
        PFEM_HIP(hipSetDevice(ctx->device));
        auto c = std::make_unique<pfem_case>();
        c->ctx = ctx;
        c->d = d;
        c->dim = d.dim;
        PFEM_HIP(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
        PFEM_HIP(hipHostMalloc((void**)&c->h_scal, 2 * sizeof(KScal), hipHostMallocDefault));
        PFEM_HIP(hipHostMalloc((void**)&c->h_out, 4 * sizeof(double), hipHostMallocDefault));
        *out = c.release();
        return PFEM_OK;
    });
}

int pfem_case_destroy(pfem_case* c) {
    return guarded([&] {
        if (!c) return PFEM_OK;
        (void)hipSetDevice(c->ctx->device);
        (void)hipStreamSynchronize(c->stream);
        delete c;
        return PFEM_OK;
    });
}

int pfem_case_set_mesh(pfem_case* c, int32_t N, int32_t E, const int32_t* conn, const uint8_t* flags,
                       const double* pos) {
    return guarded([&] {
        require_case(c);
        PFEM_REQUIRE(N > 0 && E > 0 && conn && flags && pos, PFEM_E_INVALID, "invalid mesh arguments");
        PFEM_HIP(hipSetDevice(c->ctx->device));
        const int dim = c->dim, ne = dim + 1;
        c->has_mesh = false;
        c->N = N;
        c->E = E;
        c->n = (dim + 1) * N;
        DevBuf d_conn, d_flags;
        d_conn.reserve(sizeof(int) * (size_t)E * ne);
        d_flags.reserve((size_t)N);
        PFEM_HIP(hipMemcpyAsync(d_conn.p, conn, sizeof(int) * (size_t)E * ne, hipMemcpyHostToDevice, c->stream));
        PFEM_HIP(hipMemcpyAsync(d_flags.p, flags, (size_t)N, hipMemcpyHostToDevice, c->stream));
        c->pos.reserve(sizeof(double) * (size_t)dim * N);
        c->pos_save.reserve(sizeof(double) * (size_t)dim * N);
        c->dir_vals.reserve(sizeof(double) * (size_t)dim * N);
        PFEM_HIP(hipMemcpyAsync(c->pos.p, pos, sizeof(double) * (size_t)dim * N, hipMemcpyHostToDevice, c->stream));
        PFEM_HIP(hipMemsetAsync(c->dir_vals.p, 0, sizeof(double) * (size_t)dim * N, c->stream));
        c->partials.reserve(sizeof(double) * pfem::kMaxCaseBlocks * pfem::kPartStride);
        c->counters.reserve(sizeof(unsigned) * 8);
        c->scal.reserve(sizeof(KScal) * 2);
        c->out.reserve(sizeof(double) * 4);
        PFEM_HIP(hipMemsetAsync(c->counters.p, 0, sizeof(unsigned) * 8, c->stream));
        PFEM_HIP(hipMemsetAsync(c->scal.p, 0, sizeof(KScal) * 2, c->stream));
        if (c->d.kind == PFEM_KIND_PSPG) {
            pfem::build_pattern(dim, N, E, d_conn.as<int>(), d_flags.as<uint8_t>(), c->stream, c->pat);
            const size_t n = (size_t)c->n, nnz = (size_t)c->pat.nnz;
            c->q_prev.reserve(sizeof(double) * n);
            c->x.reserve(sizeof(double) * n);
            c->q_itprev.reserve(sizeof(double) * n);
            c->vals.reserve(sizeof(double) * nnz);
            c->b.reserve(sizeof(double) * n);
            c->binv.reserve(sizeof(double) * (size_t)(dim + 1) * (dim + 1) * N);
            c->kv64.reserve(sizeof(double) * n * 8);
            PFEM_HIP(hipMemsetAsync(c->q_prev.p, 0, sizeof(double) * n, c->stream));
            PFEM_HIP(hipMemsetAsync(c->x.p, 0, sizeof(double) * n, c->stream));
            if (c->d.precision == PFEM_PREC_MIXED) {
                c->vals32.reserve(sizeof(float) * nnz);
                c->binv32.reserve(sizeof(float) * (size_t)(dim + 1) * (dim + 1) * N);
                c->kv32.reserve(sizeof(float) * n * 10);
            }
            c->lanes = pick_lanes((double)c->pat.graph_nnz / (double)N);
            fill_views(c);
        } else {
            pfem::build_mesh_basic(dim, N, E, d_conn.as<int>(), d_flags.as<uint8_t>(), c->stream, c->pat);
            const size_t ss = wc_state_size(c);
            c->wstate.reserve(ss * (size_t)(2 * dim + 2) * N);
            PFEM_HIP(hipMemsetAsync(c->wstate.p, 0, ss * (size_t)(2 * dim + 2) * N, c->stream));
            c->wacc.reserve(sizeof(double) * (size_t)(3 + dim) * N);
            c->wdt2.reserve(sizeof(unsigned long long));
            fill_wc_views(c);
        }
        PFEM_HIP(hipStreamSynchronize(c->stream));
        c->has_mesh = true;
        return PFEM_OK;
    });
}

int pfem_case_set_state(pfem_case* c, const double* q) {
    return guarded([&] {
        require_mesh(c);
        PFEM_REQUIRE(q != nullptr, PFEM_E_INVALID, "null state");
        PFEM_HIP(hipSetDevice(c->ctx->device));
        if (c->d.kind == PFEM_KIND_PSPG) {
            PFEM_HIP(hipMemcpyAsync(c->q_prev.p, q, sizeof(double) * (size_t)c->n, hipMemcpyHostToDevice, c->stream));
            PFEM_HIP(hipMemcpyAsync(c->x.p, q, sizeof(double) * (size_t)c->n, hipMemcpyHostToDevice, c->stream));
        } else {
            const size_t cnt = (size_t)(2 * c->dim + 2) * c->N;
            if (c->d.wc_state == PFEM_STATE_FP32) {
                std::vector<float> h(q, q + cnt);
                PFEM_HIP(hipMemcpyAsync(c->wstate.p, h.data(), sizeof(float) * cnt, hipMemcpyHostToDevice, c->stream));
                PFEM_HIP(hipStreamSynchronize(c->stream));
            } else {
                PFEM_HIP(hipMemcpyAsync(c->wstate.p, q, sizeof(double) * cnt, hipMemcpyHostToDevice, c->stream));
            }
        }
        PFEM_HIP(hipStreamSynchronize(c->stream));
        return PFEM_OK;
    });
}

int pfem_case_get_state(pfem_case* c, double* q) {
    return guarded([&] {
        require_mesh(c);
        PFEM_REQUIRE(q != nullptr, PFEM_E_INVALID, "null state");
        PFEM_HIP(hipSetDevice(c->ctx->device));
        if (c->d.kind == PFEM_KIND_PSPG) {
            PFEM_HIP(hipMemcpyAsync(q, c->q_prev.p, sizeof(double) * (size_t)c->n, hipMemcpyDeviceToHost, c->stream));
        } else {
            const size_t cnt = (size_t)(2 * c->dim + 2) * c->N;
            if (c->d.wc_state == PFEM_STATE_FP32) {
                std::vector<float> h(cnt);
                PFEM_HIP(hipMemcpyAsync(h.data(), c->wstate.p, sizeof(float) * cnt, hipMemcpyDeviceToHost, c->stream));
                PFEM_HIP(hipStreamSynchronize(c->stream));
                std::copy(h.begin(), h.end(), q);
            } else {
                PFEM_HIP(hipMemcpyAsync(q, c->wstate.p, sizeof(double) * cnt, hipMemcpyDeviceToHost, c->stream));
            }
        }
        PFEM_HIP(hipStreamSynchronize(c->stream));
        return PFEM_OK;
    });
}

int pfem_case_set_positions(pfem_case* c, const double* pos) {
    return guarded([&] {
        require_mesh(c);
        PFEM_REQUIRE(pos != nullptr, PFEM_E_INVALID, "null positions");
        PFEM_HIP(hipMemcpyAsync(c->pos.p, pos, sizeof(double) * (size_t)c->dim * c->N, hipMemcpyHostToDevice,
                                c->stream));
        PFEM_HIP(hipStreamSynchronize(c->stream));
        return PFEM_OK;
    });
}

int pfem_case_get_positions(pfem_case* c, double* pos) {
    return guarded([&] {
        require_mesh(c);
        PFEM_REQUIRE(pos != nullptr, PFEM_E_INVALID, "null positions");
        PFEM_HIP(hipMemcpyAsync(pos, c->pos.p, sizeof(double) * (size_t)c->dim * c->N, hipMemcpyDeviceToHost,
                                c->stream));
        PFEM_HIP(hipStreamSynchronize(c->stream));
        return PFEM_OK;
    });
}

int pfem_case_set_dirichlet(pfem_case* c, const double* vals) {
    return guarded([&] {
        require_mesh(c);
        PFEM_REQUIRE(vals != nullptr, PFEM_E_INVALID, "null Dirichlet values");
        PFEM_HIP(hipMemcpyAsync(c->dir_vals.p, vals, sizeof(double) * (size_t)c->dim * c->N, hipMemcpyHostToDevice,
                                c->stream));
        PFEM_HIP(hipStreamSynchronize(c->stream));
        return PFEM_OK;
    });
}

int pfem_case_sync(pfem_case* c) {
    return guarded([&] {
        require_case(c);
        PFEM_HIP(hipStreamSynchronize(c->stream));
        return PFEM_OK;
    });
}

int pfem_case_footprint(pfem_case* c, int64_t* bytes) {
    return guarded([&] {
        require_case(c);
        PFEM_REQUIRE(bytes != nullptr, PFEM_E_INVALID, "null output");
        size_t b = 0;
        for (DevBuf* p : {&c->pos, &c->pos_save, &c->dir_vals, &c->q_prev, &c->x, &c->q_itprev, &c->vals, &c->b,
                          &c->vals32, &c->binv, &c->binv32, &c->kv64, &c->kv32, &c->partials, &c->counters, &c->scal,
                          &c->out, &c->wstate, &c->wacc, &c->wdt2, &c->pat.conn, &c->pat.flags, &c->pat.nbr_ptr,
                          &c->pat.nbr_idx, &c->pat.node_info, &c->pat.ranks, &c->pat.row_ptr, &c->pat.col_idx,
                          &c->pat.diag_pos, &c->pat.pei_ptr, &c->pat.pei})
            b += p->bytes;
        *bytes = (int64_t)b;
        return PFEM_OK;
    });
}

int pfem_pspg_assemble(pfem_case* c, double dt, const double* q_prev, const double* q_iter) {
    return guarded([&] {
        require_mesh(c);
        PFEM_REQUIRE(c->d.kind == PFEM_KIND_PSPG, PFEM_E_STATE, "not a PSPG case");
        PFEM_REQUIRE(dt > 0, PFEM_E_INVALID, "dt must be > 0");
        PFEM_HIP(hipSetDevice(c->ctx->device));
        const size_t nb = sizeof(double) * (size_t)c->n;
        if (q_prev) PFEM_HIP(hipMemcpyAsync(c->q_prev.p, q_prev, nb, hipMemcpyHostToDevice, c->stream));
        if (q_iter) PFEM_HIP(hipMemcpyAsync(c->x.p, q_iter, nb, hipMemcpyHostToDevice, c->stream));
        c->pv.dt = dt;
        c->picv.dt = dt;
        pfem::Runner& R = case_runner(c);
        R.set_active({c});
        pfem::assemble_active(R);
        PFEM_HIP(hipStreamSynchronize(c->stream));
        return PFEM_OK;
    });
}

int pfem_pspg_solve(pfem_case* c, const double* x0, double* x_out, pfem_solve_stats* stats) {
    return guarded([&] {
        require_mesh(c);
        PFEM_REQUIRE(c->d.kind == PFEM_KIND_PSPG, PFEM_E_STATE, "not a PSPG case");
        PFEM_HIP(hipSetDevice(c->ctx->device));
        const size_t nb = sizeof(double) * (size_t)c->n;
        if (x0) PFEM_HIP(hipMemcpyAsync(c->x.p, x0, nb, hipMemcpyHostToDevice, c->stream));
        pfem::Runner& R = case_runner(c);
        R.set_active({c});
        std::vector<pfem_solve_stats> st;
        pfem::solve_active(R, {c}, st);
        if (x_out) PFEM_HIP(hipMemcpyAsync(x_out, c->x.p, nb, hipMemcpyDeviceToHost, c->stream));
        PFEM_HIP(hipStreamSynchronize(c->stream));
        if (stats) *stats = st[0];
        return st[0].status;
    });
}

int pfem_get_csr_sizes(pfem_case* c, int64_t* n, int64_t* nnz) {
    return guarded([&] {
        require_mesh(c);
        PFEM_REQUIRE(c->d.kind == PFEM_KIND_PSPG, PFEM_E_STATE, "not a PSPG case");
        if (n) *n = c->n;
        if (nnz) *nnz = c->pat.nnz;
        return PFEM_OK;
    });
}

int pfem_get_csr(pfem_case* c, int32_t* row_ptr, int32_t* col_idx, double* vals) {
    return guarded([&] {
        require_mesh(c);
        PFEM_REQUIRE(c->d.kind == PFEM_KIND_PSPG, PFEM_E_STATE, "not a PSPG case");
        if (row_ptr)
            PFEM_HIP(hipMemcpyAsync(row_ptr, c->pat.row_ptr.p, sizeof(int) * (size_t)(c->n + 1), hipMemcpyDeviceToHost,
                                    c->stream));
        if (col_idx)
            PFEM_HIP(hipMemcpyAsync(col_idx, c->pat.col_idx.p, sizeof(int) * (size_t)c->pat.nnz, hipMemcpyDeviceToHost,
                                    c->stream));
        if (vals)
            PFEM_HIP(hipMemcpyAsync(vals, c->vals.p, sizeof(double) * (size_t)c->pat.nnz, hipMemcpyDeviceToHost,
                                    c->stream));
        PFEM_HIP(hipStreamSynchronize(c->stream));
        return PFEM_OK;
    });
}

int pfem_get_rhs(pfem_case* c, double* b) {
    return guarded([&] {
        require_mesh(c);
        PFEM_REQUIRE(c->d.kind == PFEM_KIND_PSPG && b, PFEM_E_INVALID, "not a PSPG case or null output");
        PFEM_HIP(hipMemcpyAsync(b, c->b.p, sizeof(double) * (size_t)c->n, hipMemcpyDeviceToHost, c->stream));
        PFEM_HIP(hipStreamSynchronize(c->stream));
        return PFEM_OK;
    });
}

int pfem_spmv(pfem_case* c, const double* x, double* y) {
    return guarded([&] {
        require_mesh(c);
        PFEM_REQUIRE(c->d.kind == PFEM_KIND_PSPG && x && y, PFEM_E_INVALID, "not a PSPG case or null vector");
        const size_t nb = sizeof(double) * (size_t)c->n;
        double* dx = c->kv64.as<double>();
        double* dy = dx + c->n;
        PFEM_HIP(hipMemcpyAsync(dx, x, nb, hipMemcpyHostToDevice, c->stream));
        pfem::launch_spmv<double>(c->lanes, node_mat<double>(c->pat, c->dim, c->vals.as<double>()), dx, dy,
                                  grid_for_rows(c->N), c->stream);
        PFEM_HIP(hipGetLastError());
        PFEM_HIP(hipMemcpyAsync(y, dy, nb, hipMemcpyDeviceToHost, c->stream));
        PFEM_HIP(hipStreamSynchronize(c->stream));
        return PFEM_OK;
    });
}

int pfem_pspg_step(pfem_case* c, double dt, pfem_step_stats* stats) {
    return guarded([&] {
        require_mesh(c);
        PFEM_REQUIRE(c->d.kind == PFEM_KIND_PSPG, PFEM_E_STATE, "not a PSPG case");
        PFEM_REQUIRE(dt > 0, PFEM_E_INVALID, "dt must be > 0");
        PFEM_HIP(hipSetDevice(c->ctx->device));
        pfem::Runner& R = case_runner(c);
        std::vector<pfem_step_stats> st;
        pfem::pspg_step_set(R, {c}, {dt}, st);
        if (stats) *stats = st[0];
        return st[0].status;
    });
}

int pfem_wc_step(pfem_case* c, double dt, pfem_step_stats* stats) {
    return guarded([&] {
        require_mesh(c);
        PFEM_REQUIRE(c->d.kind == PFEM_KIND_WC, PFEM_E_STATE, "not a WC case");
        PFEM_REQUIRE(dt > 0, PFEM_E_INVALID, "dt must be > 0");
        PFEM_HIP(hipSetDevice(c->ctx->device));
        c->wc64.dt = c->wc32.dt = dt;
        pfem::Runner& R = case_runner(c);
        if (c->d.wc_state == PFEM_STATE_FP32) pfem::wc_step_group<float>(R, {c});
        else pfem::wc_step_group<double>(R, {c});
        PFEM_HIP(hipStreamSynchronize(c->stream));
        if (stats) {
            *stats = pfem_step_stats{};
            stats->assemblies = 2;
        }
        return PFEM_OK;
    });
}

int pfem_wc_next_dt(pfem_case* c, double security_coeff, double max_dt, double* dt_out) {
    return guarded([&] {
        require_mesh(c);
        PFEM_REQUIRE(c->d.kind == PFEM_KIND_WC && dt_out, PFEM_E_INVALID, "not a WC case or null output");
        PFEM_HIP(hipSetDevice(c->ctx->device));
        c->wc64.sc2 = c->wc32.sc2 = security_coeff * security_coeff;
        pfem::Runner& R = case_runner(c);
        R.set_active({c});
        const unsigned long long init = 0x7fefffffffffffffULL;  // DBL_MAX
        PFEM_HIP(hipMemcpyAsync(c->wdt2.p, &init, sizeof(init), hipMemcpyHostToDevice, c->stream));
        if (c->d.wc_state == PFEM_STATE_FP32)
            pfem::launch_wc_cfl<float>(R.v_wc32.as<pfem::WcView<float>>(), R.te(0), R.nt_elem[0], R.te(1),
                                       R.nt_elem[1], c->stream);
        else
            pfem::launch_wc_cfl<double>(R.v_wc64.as<pfem::WcView<double>>(), R.te(0), R.nt_elem[0], R.te(1),
                                        R.nt_elem[1], c->stream);
        PFEM_HIP(hipGetLastError());
        unsigned long long bits = 0;
        PFEM_HIP(hipMemcpyAsync(&bits, c->wdt2.p, sizeof(bits), hipMemcpyDeviceToHost, c->stream));
        PFEM_HIP(hipStreamSynchronize(c->stream));
        double dt2;
        std::memcpy(&dt2, &bits, sizeof(dt2));
        const double dt = std::min(std::sqrt(dt2), max_dt);
        PFEM_REQUIRE(!std::isnan(dt), PFEM_NAN, "NaN time step");
        *dt_out = dt;
        return PFEM_OK;
    });
}

int pfem_profile_spmv(pfem_case* c, int32_t enable, double* out4) {
    return guarded([&] {
        require_case(c);
        pfem::Runner& R = case_runner(c);
        if (out4) {
            for (int k = 0; k < 2; ++k) {
                out4[2 * k] = R.prof_count[k] ? R.prof_ms[k] / (double)R.prof_count[k] : 0.0;
                out4[2 * k + 1] = (double)R.prof_count[k];
            }
        }
        R.enable_profiling(enable != 0);
        return PFEM_OK;
    });
}

int pfem_bench_kernels(pfem_case* c, int32_t reps, double* out8) {
    return guarded([&] {
        require_mesh(c);
        PFEM_REQUIRE(c->d.kind == PFEM_KIND_PSPG && out8 && reps > 0, PFEM_E_INVALID, "invalid arguments");
        PFEM_HIP(hipSetDevice(c->ctx->device));
        hipEvent_t e0, e1;
        PFEM_HIP(hipEventCreateWithFlags(&e0, hipEventDisableSystemFence));
        PFEM_HIP(hipEventCreateWithFlags(&e1, hipEventDisableSystemFence));
        auto timed = [&](auto&& launch) {
            launch();  // warm-up
            PFEM_HIP(hipEventRecord(e0, c->stream));
            for (int r = 0; r < reps; ++r) launch();
            PFEM_HIP(hipEventRecord(e1, c->stream));
            PFEM_HIP(hipEventSynchronize(e1));
            PFEM_HIP(hipGetLastError());
            return 1e3 * (double)pfem::ev_ms(e0, e1) / reps;  // microseconds per launch
        };
        const int n = c->n;
        double* w = c->kv64.as<double>();
        for (int k = 0; k < 8; ++k) out8[k] = 0;
        out8[0] = timed([&] {
            pfem::launch_spmv<double>(c->lanes, node_mat<double>(c->pat, c->dim, c->vals.as<double>()),
                                      c->q_prev.as<double>(), w + n, grid_for_rows(c->N), c->stream);
        });
        if (c->d.precision == PFEM_PREC_MIXED) {
            float* u = c->kv32.as<float>();
            out8[1] = timed([&] {
                pfem::launch_spmv<float>(c->lanes, node_mat<float>(c->pat, c->dim, c->vals32.as<float>()), u, u + n,
                                         grid_for_rows(c->N), c->stream);
            });
        }
        const long long nnz = c->pat.nnz;
        out8[2] = timed([&] { pfem::launch_stream_read(c->vals.as<double>(), nnz, w, 2048, c->stream); });
        const long long ncopy = std::min<long long>(nnz, 4LL * n);
        out8[3] = timed([&] { pfem::launch_copy(c->vals.as<double>(), w + 4LL * n, ncopy, 2048, c->stream); });
        out8[4] = (double)ncopy * 16.0;
        out8[5] = (double)nnz;
        out8[6] = (double)c->pat.graph_nnz;
        out8[7] = (double)c->lanes;
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        return PFEM_OK;
    });
}

int pfem_probe_xcd(pfem_ctx* ctx, int32_t nblocks, int32_t* out) {
    return guarded([&] {
        PFEM_REQUIRE(ctx && out && nblocks > 0 && nblocks <= (1 << 20), PFEM_E_INVALID, "invalid arguments");
        PFEM_HIP(hipSetDevice(ctx->device));
        DevBuf d;
        d.reserve(sizeof(int) * (size_t)nblocks);
        pfem::launch_xcd_probe(d.as<int>(), nblocks, ctx->stream);
        PFEM_HIP(hipGetLastError());
        PFEM_HIP(hipMemcpyAsync(out, d.p, sizeof(int) * (size_t)nblocks, hipMemcpyDeviceToHost, ctx->stream));
        PFEM_HIP(hipStreamSynchronize(ctx->stream));
        return PFEM_OK;
    });
}

int pfem_batch_step(pfem_ctx* ctx, int32_t ncases, pfem_case* const* cases, const double* dts, int32_t* status,
                    pfem_step_stats* stats) {
    return guarded([&] {
        PFEM_REQUIRE(ctx && ncases >= 0 && (ncases == 0 || (cases && dts)), PFEM_E_INVALID, "invalid batch");
        std::lock_guard<std::mutex> lock(ctx->mu);
        PFEM_HIP(hipSetDevice(ctx->device));
        std::vector<pfem_case*> pspg, wc64, wc32;
        std::vector<double> pdts;
        std::vector<int> pidx;
        for (int i = 0; i < ncases; ++i) {
            pfem_case* c = cases[i];
            require_mesh(c);
            PFEM_REQUIRE(c->ctx == ctx, PFEM_E_INVALID, "case belongs to another context");
            PFEM_REQUIRE(dts[i] > 0, PFEM_E_INVALID, "dt must be > 0");
            PFEM_HIP(hipStreamSynchronize(c->stream));
            if (c->d.kind == PFEM_KIND_PSPG) {
                pspg.push_back(c);
                pdts.push_back(dts[i]);
                pidx.push_back(i);
            } else {
                c->wc64.dt = c->wc32.dt = dts[i];
                (c->d.wc_state == PFEM_STATE_FP32 ? wc32 : wc64).push_back(c);
            }
        }
        pfem::Runner& R = *ctx->runner;
        pfem::wc_step_group<double>(R, wc64);
        pfem::wc_step_group<float>(R, wc32);
        std::vector<pfem_step_stats> st;
        if (!pspg.empty()) pfem::pspg_step_set(R, pspg, pdts, st);
        PFEM_HIP(hipStreamSynchronize(R.s));
        for (int i = 0; i < ncases; ++i) {
            if (status) status[i] = PFEM_OK;
            if (stats) stats[i] = pfem_step_stats{};
        }
        for (size_t j = 0; j < pspg.size(); ++j) {
            if (status) status[pidx[j]] = st[j].status;
            if (stats) stats[pidx[j]] = st[j];
        }
        return PFEM_OK;
    });
}

}  // extern "C"
