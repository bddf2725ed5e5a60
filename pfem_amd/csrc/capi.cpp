// C-ABI of include/pfem_gpu.h over the tick engine (runtime.hpp): contexts, cases, the
// direct per-case calls (a one-slot engine per case), pfem_batch_step (one engine per
// context) and the serving loop (an engine with a workspace arena on its own thread).
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <limits>
#include <new>
#include <string>
#include <vector>

#include "fracstep.hpp"
#include "heat.hpp"
#include "mesher.hpp"
#include "runtime.hpp"

using pfem::DevBuf;
using pfem::Req;

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

int grid_for_rows(int n) { return pfem::row_blocks(n); }

// Heat case: the scalar CSR over the node graph (full rows hold the node's sorted graph
// entries, Dirichlet and free rows their diagonal only) and the step's device arrays.
void setup_heat(pfem_case* c) {
    const int N = c->N, E = c->E;
    const pfem::Pattern& P = c->pat;
    std::vector<int> ptr((size_t)N + 1), idx((size_t)P.graph_nnz);
    std::vector<uint8_t> fl((size_t)N);
    PFEM_HIP(hipMemcpyAsync(ptr.data(), P.nbr_ptr.p, sizeof(int) * ptr.size(), hipMemcpyDeviceToHost, c->stream));
    PFEM_HIP(hipMemcpyAsync(idx.data(), P.nbr_idx.p, sizeof(int) * idx.size(), hipMemcpyDeviceToHost, c->stream));
    PFEM_HIP(hipMemcpyAsync(fl.data(), P.flags.p, fl.size(), hipMemcpyDeviceToHost, c->stream));
    PFEM_HIP(hipStreamSynchronize(c->stream));
    auto& H = c->heat;
    H.dbit = c->d.kind == PFEM_KIND_BOUSSINESQ ? pfem::NODE_DIRICHLET_T : pfem::NODE_DIRICHLET;
    std::vector<int> rp((size_t)N + 1, 0), col, dpos((size_t)N);
    col.reserve(idx.size());
    for (int i = 0; i < N; ++i) {
        const bool full = !(fl[i] & (H.dbit | pfem::NODE_FREE));
        if (full) {
            for (int q = ptr[i]; q < ptr[i + 1]; ++q) {
                if (idx[q] == i) dpos[i] = (int)col.size();
                col.push_back(idx[q]);
            }
        } else {
            dpos[i] = (int)col.size();
            col.push_back(i);
        }
        rp[i + 1] = (int)col.size();
    }
    H.nnz = (long long)col.size();
    auto up = [&](pfem::DevBuf& d, const void* src, size_t bytes) {
        d.reserve(bytes);
        PFEM_HIP(hipMemcpyAsync(d.p, src, bytes, hipMemcpyHostToDevice, c->stream));
    };
    up(H.row_ptr, rp.data(), sizeof(int) * rp.size());
    up(H.col, col.data(), sizeof(int) * col.size());
    up(H.diag_pos, dpos.data(), sizeof(int) * dpos.size());
    const int nb = pfem::heat_blocks(N);
    H.vals.reserve(sizeof(double) * col.size());
    for (pfem::DevBuf* b : {&H.b, &H.x, &H.r, &H.z, &H.p, &H.q, &H.dinv}) b->reserve(sizeof(double) * (size_t)N);
    H.erec.reserve(sizeof(double) * 16 * (size_t)E);
    H.part.reserve(sizeof(double) * 5 * (size_t)nb);
    H.scal.reserve(sizeof(double) * 2);
    H.T.reserve(sizeof(double) * (size_t)N);
    PFEM_HIP(hipMemsetAsync(H.T.p, 0, sizeof(double) * (size_t)N, c->stream));
    H.Tbc.reserve(sizeof(double) * (size_t)N);
    PFEM_HIP(hipMemsetAsync(H.Tbc.p, 0, sizeof(double) * (size_t)N, c->stream));
    // the facets of a previous mesh no longer apply
    H.F = 0;
    H.has_q = false;
    std::vector<int> zeros((size_t)N + 1, 0);
    up(H.nf_ptr, zeros.data(), sizeof(int) * zeros.size());
}

pfem::HeatView heat_view(pfem_case* c, double dt) {
    auto& H = c->heat;
    pfem::HeatView V;
    V.dim = c->dim;
    V.N = c->N;
    V.E = c->E;
    V.F = H.F;
    V.nb = pfem::heat_blocks(c->N);
    V.conn = c->pat.conn.as<int>();
    V.flags = c->pat.flags.as<uint8_t>();
    V.pos = c->pos.as<double>();
    V.nbr_ptr = c->pat.nbr_ptr.as<int>();
    V.nbr_idx = c->pat.nbr_idx.as<int>();
    V.pei_ptr = c->pat.pei_ptr.as<int>();
    V.pei = c->pat.pei.as<unsigned>();
    V.row_ptr = H.row_ptr.as<int>();
    V.col = H.col.as<int>();
    V.diag_pos = H.diag_pos.as<int>();
    V.vals = H.vals.as<double>();
    V.b = H.b.as<double>();
    V.erec = H.erec.as<double>();
    V.dbit = H.dbit;
    V.Tprev = H.T.as<double>();
    V.Tbc = H.Tbc.as<double>();
    V.facets = H.facets.as<int>();
    V.fout = H.fout.as<int>();
    V.fkind = H.fkind.as<uint8_t>();
    V.qgp = H.has_q ? H.qgp.as<double>() : nullptr;
    V.fterm = H.fterm.as<double>();
    V.nf_ptr = H.nf_ptr.as<int>();
    V.nf_list = H.nf_list.as<int>();
    V.rho = c->d.rho;
    V.cv = c->d.cv;
    V.k = c->d.k_cond;
    V.h = c->d.h_conv;
    V.Tinf = c->d.t_inf;
    V.epsRad = c->d.eps_rad;
    V.dt = dt;
    V.x = H.x.as<double>();
    V.r = H.r.as<double>();
    V.z = H.z.as<double>();
    V.p = H.p.as<double>();
    V.q = H.q.as<double>();
    V.dinv = H.dinv.as<double>();
    V.part = H.part.as<double>();
    V.scal = H.scal.as<double>();
    return V;
}


// Fractional-step case: the three graph matrices, element records, right-hand sides, the
// Picard iterate and the CG vectors.
void setup_frac(pfem_case* c) {
    const int D = c->dim, N = c->N, E = c->E;
    const size_t G = (size_t)c->pat.graph_nnz, vN = (size_t)D * N;
    auto& F = c->frac;
    F.erec.reserve(sizeof(double) * 16 * (size_t)E);
    F.mk.reserve(sizeof(double) * G * D * D);
    F.ms.reserve(sizeof(double) * G);
    F.lv.reserve(sizeof(double) * G);
    for (DevBuf* b : {&F.dinv_mk, &F.dinv_m, &F.bv, &F.vt, &F.dv, &F.v_it, &F.v_itp, &F.pos_save, &F.r, &F.z, &F.p,
                      &F.q})
        b->reserve(sizeof(double) * vN);
    for (DevBuf* b : {&F.dinv_l, &F.bp, &F.pn, &F.p_itp}) b->reserve(sizeof(double) * (size_t)N);
    F.res_part.reserve(sizeof(double) * 4 * (size_t)pfem::frac_blocks(N));
    F.part.reserve(sizeof(double) * 5 * (size_t)pfem::frac_blocks(D * N));
    F.scal.reserve(sizeof(double) * 2);
    const size_t nq = sizeof(double) * (size_t)(D + 1) * N;
    c->q_prev.reserve(nq);
    PFEM_HIP(hipMemsetAsync(c->q_prev.p, 0, nq, c->stream));
}

pfem::FracView frac_view(pfem_case* c, double dt) {
    auto& F = c->frac;
    const int D = c->dim, N = c->N;
    pfem::FracView V;
    V.dim = D;
    V.N = N;
    V.E = c->E;
    V.nb = pfem::frac_blocks(N);
    V.conn = c->pat.conn.as<int>();
    V.flags = c->pat.flags.as<uint8_t>();
    V.pos = c->pos.as<double>();
    V.pos_save = F.pos_save.as<double>();
    V.nbr_ptr = c->pat.nbr_ptr.as<int>();
    V.nbr_idx = c->pat.nbr_idx.as<int>();
    V.pei_ptr = c->pat.pei_ptr.as<int>();
    V.pei = c->pat.pei.as<unsigned>();
    V.erec = F.erec.as<double>();
    V.mk = F.mk.as<double>();
    V.ms = F.ms.as<double>();
    V.lv = F.lv.as<double>();
    V.dinv_mk = F.dinv_mk.as<double>();
    V.dinv_m = F.dinv_m.as<double>();
    V.dinv_l = F.dinv_l.as<double>();
    V.bv = F.bv.as<double>();
    V.bp = F.bp.as<double>();
    V.v_prev = c->q_prev.as<double>();
    V.p_prev = c->q_prev.as<double>() + (size_t)D * N;
    V.vbc = c->dir_vals.as<double>();
    V.vt = F.vt.as<double>();
    V.pn = F.pn.as<double>();
    V.dv = F.dv.as<double>();
    V.v_it = F.v_it.as<double>();
    V.v_itp = F.v_itp.as<double>();
    V.p_itp = F.p_itp.as<double>();
    V.res_part = F.res_part.as<double>();
    V.rho = c->d.rho;
    V.mu = c->d.mu;
    V.dt = dt;
    V.gamma_fs = c->d.gamma_fs;
    V.problem = c->d.problem;
    V.tau0 = c->d.tau0;
    V.mreg = c->d.m_reg;
    for (int k = 0; k < 3; ++k) V.g[k] = c->d.body_force[k];
    return V;
}

// op: 0 L (N rows), 1 M per velocity component, 2 M/dt + K (dim N rows)
pfem::CgView frac_cg(pfem_case* c, int op, const double* vals, const double* b, const double* dinv, double* x) {
    auto& F = c->frac;
    pfem::CgView C;
    C.op = op;
    C.dim = c->dim;
    C.N = c->N;
    C.n = op == 0 ? c->N : c->dim * c->N;
    C.nb = pfem::frac_blocks(C.n);
    C.nbr_ptr = c->pat.nbr_ptr.as<int>();
    C.nbr_idx = c->pat.nbr_idx.as<int>();
    C.vals = vals;
    C.b = b;
    C.dinv = dinv;
    C.x = x;
    C.r = F.r.as<double>();
    C.z = F.z.as<double>();
    C.p = F.p.as<double>();
    C.q = F.q.as<double>();
    C.part = F.part.as<double>();
    C.scal = F.scal.as<double>();
    return C;
}

// Conjugate gradient from a zero guess until ||b - A x||^2 <= rtol^2 ||b||^2 (r.r checked by
// the host every 8 iterations; a chunk may run a few iterations past convergence, which CG
// keeps reducing).  Returns PFEM_OK, PFEM_NOT_CONVERGED or PFEM_NAN.
int frac_cg_solve(pfem_case* c, const pfem::CgView& C, int* iters, double* relres) {
    PFEM_HIP(hipMemsetAsync(C.x, 0, sizeof(double) * (size_t)C.n, c->stream));
    pfem::launch_cg2_init(C, c->stream);
    double sc[2] = {0, 0};
    PFEM_HIP(hipMemcpyAsync(sc, C.scal, sizeof sc, hipMemcpyDeviceToHost, c->stream));
    PFEM_HIP(hipStreamSynchronize(c->stream));
    const double bb = sc[1], thr = c->d.krylov_rtol * c->d.krylov_rtol * bb;
    double rr = sc[0];
    int it = 0;
    while (std::isfinite(rr) && rr > thr && it < c->d.krylov_max_iter) {
        const int chunk = std::min(8, c->d.krylov_max_iter - it);
        for (int k = 0; k < chunk; ++k, ++it) pfem::launch_cg2_iteration(C, it & 1, c->stream);
        PFEM_HIP(hipMemcpyAsync(sc, C.scal, sizeof sc, hipMemcpyDeviceToHost, c->stream));
        PFEM_HIP(hipStreamSynchronize(c->stream));
        rr = sc[0];
    }
    PFEM_HIP(hipGetLastError());
    *iters = it;
    *relres = bb > 0 ? std::sqrt(rr / bb) : 0.0;
    if (!std::isfinite(rr)) return PFEM_NAN;
    return rr > thr ? PFEM_NOT_CONVERGED : PFEM_OK;
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

void require_case(pfem_case* c) {
    PFEM_REQUIRE(c != nullptr, PFEM_E_INVALID, "null case handle");
    PFEM_REQUIRE(c->busy.load() == 0, PFEM_E_STATE, "case is held by a serving request");
}
void require_mesh(pfem_case* c) {
    require_case(c);
    PFEM_REQUIRE(c->has_mesh, PFEM_E_STATE, "no mesh set on this case");
}
void require_pspg(pfem_case* c) {
    require_mesh(c);
    PFEM_REQUIRE(pfem::pspg_like(c), PFEM_E_STATE, "not a PSPG case");
}
void require_heat(pfem_case* c) {
    require_mesh(c);
    PFEM_REQUIRE(pfem::heat_like(c), PFEM_E_STATE, "not a heat case");
}
void require_frac(pfem_case* c) {
    require_mesh(c);
    PFEM_REQUIRE(c->d.kind == PFEM_KIND_FRACSTEP, PFEM_E_STATE, "not a fractional-step case");
}
// The fused Picard step takes the Dirichlet values once per step; a Dirichlet node that
// moves needs them re-evaluated at its new position before every re-assembly
// (MomContEquationPSPG.inl:206-214), which the callback-level entry points allow.
void require_static_dirichlet(pfem_case* c) {
    PFEM_REQUIRE(!c->moving_dirichlet, PFEM_E_UNSUPPORTED,
                 "the mesh has Dirichlet nodes that are not fixed: their values change with their positions at "
                 "every Picard re-assembly; drive the step through pfem_save_positions / pfem_pspg_assemble / "
                 "pfem_pspg_solve / pfem_update_positions / pfem_residual with pfem_case_set_dirichlet before each "
                 "assembly");
}
void require_system(pfem_case* c) {
    require_pspg(c);
    PFEM_REQUIRE(c->ws.base != nullptr, PFEM_E_STATE, "no assembled system on this case (assemble first)");
}

pfem::Engine& case_engine(pfem_case* c) {
    if (!c->engine) c->engine = std::make_unique<pfem::Engine>(c->ctx->device, c->stream, 1, nullptr);
    return *c->engine;
}

// Runs one request on the case's own engine until it completes.
int run_direct(pfem_case* c, Req& r) {
    pfem::Engine& E = case_engine(c);
    std::vector<Req*> done;
    r.c = c;
    r.t_arrival = pfem::now_ns();
    E.enqueue(&r);
    try {
        E.drain(done);
    } catch (...) {
        std::vector<Req*> d2;
        E.abort_all(PFEM_E_HIP, d2);
        throw;
    }
    return r.status;
}

size_t wc_state_size(pfem_case* c) { return c->d.wc_state == PFEM_STATE_FP32 ? sizeof(float) : sizeof(double); }

size_t persistent_bytes(pfem_case* c) {
    size_t b = 0;
    for (DevBuf* p : {&c->pos, &c->dir_vals, &c->q_prev, &c->partials, &c->counters, &c->wstate, &c->wacc, &c->wrec, &c->werec,
                      &c->pat.conn, &c->pat.flags, &c->pat.nbr_ptr, &c->pat.nbr_idx, &c->pat.node_info,
                      &c->pat.row_ptr, &c->pat.diag_pos, &c->pat.pei_ptr, &c->pat.pei, &c->pat.elem_ptr,
                      &c->pat.elem_list, &c->pat.elem_nodes, &c->pat.ell_soff, &c->pat.ell_idx})
        b += p->bytes;
    auto& H = c->heat;
    for (DevBuf* p : {&H.row_ptr, &H.col, &H.diag_pos, &H.vals, &H.b, &H.erec, &H.facets, &H.fout, &H.fkind, &H.qgp,
                      &H.fterm, &H.nf_ptr, &H.nf_list, &H.x, &H.r, &H.z, &H.p, &H.q, &H.dinv, &H.part, &H.scal, &H.T,
                      &H.Tbc})
        b += p->bytes;
    auto& F = c->frac;
    for (DevBuf* p : {&F.erec, &F.mk, &F.ms, &F.lv, &F.dinv_mk, &F.dinv_m, &F.dinv_l, &F.bv, &F.bp, &F.vt, &F.pn,
                      &F.dv, &F.v_it, &F.v_itp, &F.p_itp, &F.pos_save, &F.res_part, &F.r, &F.z, &F.p, &F.q, &F.part,
                      &F.scal})
        b += p->bytes;
    for (DevBuf* p : {&c->fst.nodes, &c->fst.elem, &c->fst.ptr, &c->fst.ent}) b += p->bytes;
    for (auto& L : c->amg.lev)
        for (DevBuf* p : {&L.nbr_ptr, &L.nbr_idx, &L.node_info, &L.self_rank, &L.agg_ptr, &L.agg_nodes, &L.fine_agg,
                          &L.gal_ptr, &L.gal_src, &L.ap_ptr, &L.ap_gptr, &L.ap_gsrc, &L.ap_soff,
                          &L.ap_eidx})
            b += p->bytes;
    return b;
}

// Multigrid for the systems whose block-Jacobi BiCGSTAB iteration count grows with the
// number of nodes across the domain: 2D cases from 2,000 nodes, 3D cases from 50,000
// (measured: below these the extra V-cycle work outweighs the saved iterations; the 3D
// 262k-node dam break steps 14.5 times per second with it against 11.7 without, and the
// one-at-a-time serving stream is fastest with 50,000 among 5k / 20k / 50k / 100k / 400k;
// in 2D, 1,000 / 2,500 / 5,000 served the stream at 77.2 / 75.3 / 73.9 steps/s one at a time
// and 224.3 / 223.6 / 220.3 batched).
// PFEM_PREC_MIXED 3D cases run the hierarchy in FP32, which halves the V-cycle's bytes:
// they take it from 5,000 nodes (the batched stream: 217.6 / 219.2 / 220.4 steps/s from
// 20k / 10k / 5k).
// PFEM_AMG_MIN_NODES_2D / _3D / _3D_MIXED override the thresholds.
bool wants_amg(const pfem_case* c) {
    if (c->d.precond == PFEM_PRECOND_JACOBI) return false;
    if (c->d.precond == PFEM_PRECOND_AMG) return true;
    auto env = [](const char* k, int dflt) {
        const char* e = std::getenv(k);
        return e ? std::atoi(e) : dflt;
    };
    static const int min2 = env("PFEM_AMG_MIN_NODES_2D", 2000), min3 = env("PFEM_AMG_MIN_NODES_3D", 50000),
                     min3m = env("PFEM_AMG_MIN_NODES_3D_MIXED", 5000);
    if (c->dim == 2) return c->N >= min2;
    return c->N >= (c->d.precision == PFEM_PREC_MIXED ? min3m : min3);
}

}  // namespace

pfem_case::~pfem_case() {
    engine.reset();
    ws.clear();
    if (stream) (void)hipStreamDestroy(stream);
}

// =========================================================================================
// serving loop
// =========================================================================================
struct pfem_server {
    pfem_ctx* ctx = nullptr;
    hipStream_t stream = nullptr;
    pfem::Arena arena;
    std::unique_ptr<pfem::Engine> eng;
    std::thread th;
    std::mutex mu;
    std::condition_variable cv, cv_done;
    std::multimap<int64_t, Req*> incoming;  // by arrival time
    std::deque<Req*> results;
    bool stop = false;
    int err = 0;
    std::string err_msg;
    double wc_sc = 0.5, wc_max_dt = 1.0;
    int64_t submitted = 0, completed = 0;
    // engine statistics, copied under the lock after every tick
    int64_t ticks = 0, peak_inflight = 0, arena_peak = 0, arena_used = 0;
};

namespace {

void server_loop(pfem_server* S) {
    (void)hipSetDevice(S->ctx->device);
    for (;;) {
        {
            std::unique_lock<std::mutex> lk(S->mu);
            const int64_t now = pfem::now_ns();
            while (!S->incoming.empty() && S->incoming.begin()->first <= now) {
                S->eng->enqueue(S->incoming.begin()->second);
                S->incoming.erase(S->incoming.begin());
            }
            if (S->eng->idle()) {
                if (S->stop && S->incoming.empty()) break;
                if (S->incoming.empty()) {
                    S->cv.wait(lk);
                } else {
                    const auto t = std::chrono::steady_clock::time_point(
                        std::chrono::nanoseconds(S->incoming.begin()->first));
                    S->cv.wait_until(lk, t);
                }
                continue;
            }
        }
        std::vector<Req*> done;
        int code = 0;
        std::string msg;
        try {
            S->eng->tick(done);
        } catch (const pfem::Error& e) {
            code = e.code;
            msg = e.what();
        } catch (const std::exception& e) {
            code = PFEM_E_HIP;
            msg = e.what();
        }
        std::lock_guard<std::mutex> lk(S->mu);
        S->ticks = S->eng->ticks;
        S->peak_inflight = S->eng->peak_inflight;
        S->arena_peak = (int64_t)S->arena.peak();
        S->arena_used = (int64_t)S->arena.used();
        if (code) {
            // a device error ends the server: fail everything in flight and queued
            S->err = code;
            S->err_msg = msg;
            S->eng->abort_all(code, done);
            for (auto& kv : S->incoming) {
                kv.second->status = code;
                kv.second->completed = true;
                done.push_back(kv.second);
            }
            S->incoming.clear();
        }
        for (Req* r : done) {
            r->c->busy.store(0);
            S->results.push_back(r);
            S->completed++;
        }
        if (!done.empty()) S->cv_done.notify_all();
        if (code) {
            S->stop = true;
            break;
        }
    }
}

}  // namespace

// =========================================================================================
// C-ABI
// =========================================================================================
extern "C" {

const char* pfem_last_error(void) { return g_err.c_str(); }
int pfem_version(void) { return 2; }
int64_t pfem_now_ns(void) { return pfem::now_ns(); }

int pfem_ctx_create(int32_t device, pfem_ctx** out) {
    return guarded([&] {
        PFEM_REQUIRE(out != nullptr, PFEM_E_INVALID, "null output pointer");
        int count = 0;
        PFEM_HIP(hipGetDeviceCount(&count));
        PFEM_REQUIRE(device >= 0 && device < count, PFEM_E_INVALID, "device index out of range");
        PFEM_HIP(hipSetDevice(device));
        pfem::init_scratch(device);
        auto ctx = std::make_unique<pfem_ctx>();
        ctx->device = device;
        PFEM_HIP(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking));
        *out = ctx.release();
        return PFEM_OK;
    });
}

int pfem_ctx_destroy(pfem_ctx* ctx) {
    return guarded([&] {
        if (!ctx) return PFEM_OK;
        (void)hipSetDevice(ctx->device);
        (void)hipStreamSynchronize(ctx->stream);
        ctx->engine.reset();
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
        PFEM_REQUIRE(d.kind == PFEM_KIND_PSPG || d.kind == PFEM_KIND_WC || d.kind == PFEM_KIND_HEAT ||
                         d.kind == PFEM_KIND_FRACSTEP || d.kind == PFEM_KIND_BOUSSINESQ,
                     PFEM_E_INVALID, "unknown case kind");
        PFEM_REQUIRE(d.gamma >= 0, PFEM_E_INVALID, "gamma must be >= 0");
        // the fractional step's pressure correction has no free-surface condition, so the only
        // flows the FST term acts on have no consistent pressure system there
        PFEM_REQUIRE(d.kind != PFEM_KIND_FRACSTEP || d.gamma < 1e-15, PFEM_E_UNSUPPORTED,
                     "surface tension is not supported by the fractional-step solver");
        PFEM_REQUIRE(d.mu >= 0, PFEM_E_INVALID, "mu must be >= 0");
        if (d.kind == PFEM_KIND_BOUSSINESQ) {
            PFEM_REQUIRE(d.cv > 0 && d.k_cond >= 0 && d.h_conv >= 0 && d.eps_rad >= 0, PFEM_E_INVALID,
                         "heat needs cv > 0, k >= 0, h >= 0, epsRad >= 0");
            PFEM_REQUIRE(d.problem == PFEM_PROBLEM_NEWTONIAN, PFEM_E_INVALID, "a Boussinesq fluid is Newtonian");
            PFEM_REQUIRE(d.dgamma_dt >= 0, PFEM_E_INVALID, "DgammaDT must be >= 0");
        }
        if (d.kind == PFEM_KIND_PSPG || d.kind == PFEM_KIND_BOUSSINESQ) {
            PFEM_REQUIRE(d.rho > 0, PFEM_E_INVALID, "rho must be > 0");
            PFEM_REQUIRE(d.residual >= 0 && d.residual <= 2, PFEM_E_INVALID, "unknown residual kind");
            PFEM_REQUIRE(d.max_iter >= 0, PFEM_E_INVALID, "maxIter must be >= 0");
            PFEM_REQUIRE(d.precision == PFEM_PREC_FP64 || d.precision == PFEM_PREC_MIXED ||
                             d.precision == PFEM_PREC_MIXED_IR,
                         PFEM_E_INVALID,
                         "unknown precision");
            PFEM_REQUIRE(d.krylov_rtol > 0 && d.krylov_max_iter > 0, PFEM_E_INVALID, "invalid Krylov settings");
            PFEM_REQUIRE(d.precond >= PFEM_PRECOND_AUTO && d.precond <= PFEM_PRECOND_AMG, PFEM_E_INVALID,
                         "unknown preconditioner");
            PFEM_REQUIRE(d.precision != PFEM_PREC_MIXED_IR || (d.inner_rtol > 0 && d.inner_max_iter > 0),
                         PFEM_E_INVALID,
                         "inner_rtol and inner_max_iter must be > 0");
            PFEM_REQUIRE(d.problem == PFEM_PROBLEM_NEWTONIAN || d.problem == PFEM_PROBLEM_BINGHAM, PFEM_E_INVALID,
                         "unknown problem");
            PFEM_REQUIRE(d.problem != PFEM_PROBLEM_BINGHAM || (d.tau0 >= 0 && d.m_reg > 0), PFEM_E_INVALID,
                         "Bingham needs tau0 >= 0 and mReg > 0");
        } else if (d.kind == PFEM_KIND_FRACSTEP) {
            PFEM_REQUIRE(d.rho > 0, PFEM_E_INVALID, "rho must be > 0");
            PFEM_REQUIRE(d.residual == PFEM_RES_U || d.residual == PFEM_RES_U_P || d.residual == PFEM_RES_AX_F,
                         PFEM_E_INVALID, "unknown residual kind");
            // MomContEquationFracStep.inl:626: "Ax-f residual currently unsupported"
            PFEM_REQUIRE(d.residual != PFEM_RES_AX_F, PFEM_E_UNSUPPORTED,
                         "the Ax_f residual is not supported by the fractional-step solver");
            PFEM_REQUIRE(d.max_iter >= 0, PFEM_E_INVALID, "maxIter must be >= 0");
            PFEM_REQUIRE(d.krylov_rtol > 0 && d.krylov_max_iter > 0, PFEM_E_INVALID, "invalid Krylov settings");
            PFEM_REQUIRE(d.gamma_fs >= 0 && d.gamma_fs <= 1, PFEM_E_INVALID, "gammaFS should be between 0 and 1");
            PFEM_REQUIRE(d.problem == PFEM_PROBLEM_NEWTONIAN || d.problem == PFEM_PROBLEM_BINGHAM, PFEM_E_INVALID,
                         "unknown problem");
            PFEM_REQUIRE(d.problem != PFEM_PROBLEM_BINGHAM || (d.tau0 >= 0 && d.m_reg > 0), PFEM_E_INVALID,
                         "Bingham needs tau0 >= 0 and mReg > 0");
        } else if (d.kind == PFEM_KIND_HEAT) {
            PFEM_REQUIRE(d.rho > 0 && d.cv > 0 && d.k_cond >= 0 && d.h_conv >= 0 && d.eps_rad >= 0, PFEM_E_INVALID,
                         "heat needs rho > 0, cv > 0, k >= 0, h >= 0, epsRad >= 0");
            PFEM_REQUIRE(d.krylov_rtol > 0 && d.krylov_max_iter > 0, PFEM_E_INVALID, "invalid Krylov settings");
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
        *out = c.release();
        return PFEM_OK;
    });
}

int pfem_case_destroy(pfem_case* c) {
    return guarded([&] {
        if (!c) return PFEM_OK;
        PFEM_REQUIRE(c->busy.load() == 0, PFEM_E_STATE, "case is held by a serving request");
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
        const int dim = c->dim, ne = dim + 1;
        // unknowns, and the pair-element incidence (E (d+1)^2 entries), are indexed in int32
        PFEM_REQUIRE((long long)ne * N <= INT32_MAX && (long long)ne * ne * E <= INT32_MAX, PFEM_E_INVALID,
                     "mesh too large: (d+1) N and (d+1)^2 E must fit 32-bit indices");
        PFEM_HIP(hipSetDevice(c->ctx->device));
        const auto t_start = std::chrono::steady_clock::now();
        auto lap = [&](int k, std::chrono::steady_clock::time_point& t0) {
            PFEM_HIP(hipStreamSynchronize(c->stream));
            const auto t1 = std::chrono::steady_clock::now();
            c->setup_s[k] = std::chrono::duration<double>(t1 - t0).count();
            t0 = t1;
        };
        for (double& t : c->setup_s) t = 0;
        c->has_mesh = false;
        c->ws.clear();
        c->cfl_valid = false;
        c->fst.F = 0;  // the facets of the previous mesh no longer apply
        c->moving_dirichlet = false;
        for (int32_t i = 0; i < N; ++i)
            if ((flags[i] & PFEM_NODE_BOUND) && (flags[i] & PFEM_NODE_DIRICHLET) && !(flags[i] & PFEM_NODE_FIXED))
                c->moving_dirichlet = true;
        c->N = N;
        c->E = E;
        c->n = (dim + 1) * N;
        // Delaunay tetrahedrisations of PFEM node clouds hold slivers (shape quality down to
        // ~1e-4; ~2 % of the elements below 0.05) that lift the spectral radius of the
        // block-Jacobi-preconditioned PSPG operator from ~2.5 to ~3.7, where the multigrid's
        // smoother damping of 0.9 amplifies instead of smoothing and BiCGSTAB diverges; 0.5
        // converges on them (107 vs 90 iterations on 136k-node meshes).  Lattice-split meshes
        // (quality >= 0.4) keep the faster 0.9.  2D Delaunay meshes converge at 0.9.
        c->min_quality = 1.0;
        c->amg_omega_cap = 1e30;
        if (dim == 3) {
            double qmin = 1.0;
            for (int32_t e = 0; e < E; ++e) {
                const int32_t* v = conn + (size_t)4 * e;
                double p[4][3];
                for (int a = 0; a < 4; ++a)
                    for (int k = 0; k < 3; ++k) p[a][k] = pos[(size_t)k * N + v[a]];
                double u[3][3], l2 = 0;
                for (int a = 0; a < 3; ++a)
                    for (int k = 0; k < 3; ++k) u[a][k] = p[a + 1][k] - p[0][k];
                for (int a = 0; a < 4; ++a)
                    for (int b = a + 1; b < 4; ++b)
                        for (int k = 0; k < 3; ++k) l2 += (p[a][k] - p[b][k]) * (p[a][k] - p[b][k]);
                const double vol = (u[0][0] * (u[1][1] * u[2][2] - u[1][2] * u[2][1]) -
                                    u[0][1] * (u[1][0] * u[2][2] - u[1][2] * u[2][0]) +
                                    u[0][2] * (u[1][0] * u[2][1] - u[1][1] * u[2][0])) / 6.0;
                const double lr = std::sqrt(l2 / 6.0);
                qmin = std::min(qmin, 6.0 * std::sqrt(2.0) * vol / (lr * lr * lr));
            }
            c->min_quality = qmin;
            if (qmin < 0.1) c->amg_omega_cap = 0.5;
        }
        DevBuf d_conn, d_flags;
        d_conn.reserve(sizeof(int) * (size_t)E * ne);
        d_flags.reserve((size_t)N);
        PFEM_HIP(hipMemcpyAsync(d_conn.p, conn, sizeof(int) * (size_t)E * ne, hipMemcpyHostToDevice, c->stream));
        PFEM_HIP(hipMemcpyAsync(d_flags.p, flags, (size_t)N, hipMemcpyHostToDevice, c->stream));
        c->pos.reserve(sizeof(double) * (size_t)dim * N);
        c->dir_vals.reserve(sizeof(double) * (size_t)dim * N);
        PFEM_HIP(hipMemcpyAsync(c->pos.p, pos, sizeof(double) * (size_t)dim * N, hipMemcpyHostToDevice, c->stream));
        PFEM_HIP(hipMemsetAsync(c->dir_vals.p, 0, sizeof(double) * (size_t)dim * N, c->stream));
        c->partials.reserve(sizeof(double) * pfem::kMaxCaseBlocks * pfem::kPartStride);
        c->counters.reserve(sizeof(unsigned) * 8);
        PFEM_HIP(hipMemsetAsync(c->counters.p, 0, sizeof(unsigned) * 8, c->stream));
        if (pfem::pspg_like(c)) {
            auto t0 = std::chrono::steady_clock::now();
            pfem::build_pattern(dim, N, E, d_conn.as<int>(), d_flags.as<uint8_t>(), c->stream, c->pat);
            lap(0, t0);
            pfem::build_ell(c->pat, c->stream);
            lap(1, t0);
            const size_t n = (size_t)c->n;
            c->q_prev.reserve(sizeof(double) * n);
            PFEM_HIP(hipMemsetAsync(c->q_prev.p, 0, sizeof(double) * n, c->stream));
            c->lanes = pfem::pick_lanes((double)c->pat.graph_nnz / (double)N);
            c->amg = pfem::AmgHierarchy{};
            const char* f32 = std::getenv("PFEM_AMG_FP32");
            c->amg_all_fp32 = f32 && std::atoi(f32) != 0;
            if (wants_amg(c)) pfem::build_amg(c->pat, dim, pfem::kAmgMaxDense / (dim + 1), c->stream, c->amg);
            lap(2, t0);
            if (c->d.kind == PFEM_KIND_BOUSSINESQ) setup_heat(c);
        } else if (c->d.kind == PFEM_KIND_HEAT) {
            pfem::build_pattern(dim, N, E, d_conn.as<int>(), d_flags.as<uint8_t>(), c->stream, c->pat);
            setup_heat(c);
        } else if (c->d.kind == PFEM_KIND_FRACSTEP) {
            pfem::build_pattern(dim, N, E, d_conn.as<int>(), d_flags.as<uint8_t>(), c->stream, c->pat);
            setup_frac(c);
        } else {
            pfem::build_mesh_basic(dim, N, E, d_conn.as<int>(), d_flags.as<uint8_t>(), c->stream, c->pat);
            const size_t ss = wc_state_size(c);
            c->wstate.reserve(ss * (size_t)(2 * dim + 2) * N);
            PFEM_HIP(hipMemsetAsync(c->wstate.p, 0, ss * (size_t)(2 * dim + 2) * N, c->stream));
            c->wacc.reserve(sizeof(double) * (size_t)(4 + dim) * N);
            c->wrec.reserve(sizeof(double) * 8 * (size_t)N);
            c->werec.reserve(sizeof(double) * 4 * (size_t)E);
            PFEM_HIP(hipMemsetAsync(c->wrec.p, 0, sizeof(double) * 8 * (size_t)N, c->stream));
        }
        PFEM_HIP(hipStreamSynchronize(c->stream));
        c->setup_s[3] = std::chrono::duration<double>(std::chrono::steady_clock::now() - t_start).count();
        c->has_mesh = true;
        return PFEM_OK;
    });
}

int pfem_case_set_state(pfem_case* c, const double* q) {
    return guarded([&] {
        require_mesh(c);
        PFEM_REQUIRE(q != nullptr, PFEM_E_INVALID, "null state");
        c->cfl_valid = false;
        PFEM_HIP(hipSetDevice(c->ctx->device));
        if (pfem::pspg_like(c) || c->d.kind == PFEM_KIND_FRACSTEP) {
            const size_t nb = sizeof(double) * (size_t)c->n;
            PFEM_HIP(hipMemcpyAsync(c->q_prev.p, q, nb, hipMemcpyHostToDevice, c->stream));
            if (c->ws.base) PFEM_HIP(hipMemcpyAsync(c->ws.x, q, nb, hipMemcpyHostToDevice, c->stream));
            if (c->d.kind == PFEM_KIND_BOUSSINESQ)  // (v, p, T): the temperatures follow
                PFEM_HIP(hipMemcpyAsync(c->heat.T.p, q + c->n, sizeof(double) * (size_t)c->N, hipMemcpyHostToDevice,
                                        c->stream));
        } else if (c->d.kind == PFEM_KIND_HEAT) {
            PFEM_HIP(hipMemcpyAsync(c->heat.T.p, q, sizeof(double) * (size_t)c->N, hipMemcpyHostToDevice, c->stream));
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
        if (pfem::pspg_like(c) || c->d.kind == PFEM_KIND_FRACSTEP) {
            PFEM_HIP(hipMemcpyAsync(q, c->q_prev.p, sizeof(double) * (size_t)c->n, hipMemcpyDeviceToHost, c->stream));
            if (c->d.kind == PFEM_KIND_BOUSSINESQ)
                PFEM_HIP(hipMemcpyAsync(q + c->n, c->heat.T.p, sizeof(double) * (size_t)c->N, hipMemcpyDeviceToHost,
                                        c->stream));
        } else if (c->d.kind == PFEM_KIND_HEAT) {
            PFEM_HIP(hipMemcpyAsync(q, c->heat.T.p, sizeof(double) * (size_t)c->N, hipMemcpyDeviceToHost, c->stream));
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
        c->cfl_valid = false;
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
        PFEM_REQUIRE(c != nullptr, PFEM_E_INVALID, "null case handle");
        PFEM_HIP(hipStreamSynchronize(c->stream));
        return PFEM_OK;
    });
}

int pfem_case_footprint(pfem_case* c, int64_t* bytes) {
    return guarded([&] {
        PFEM_REQUIRE(c != nullptr && bytes != nullptr, PFEM_E_INVALID, "null argument");
        *bytes = (int64_t)(persistent_bytes(c) + pfem::workspace_bytes(c));
        return PFEM_OK;
    });
}

int pfem_case_engine_stats(pfem_case* c, int64_t* out4) {
    return guarded([&] {
        PFEM_REQUIRE(c != nullptr && out4 != nullptr, PFEM_E_INVALID, "null argument");
        for (int k = 0; k < 4; ++k) out4[k] = 0;
        if (!c->engine) return PFEM_OK;
        long long g[2];
        c->engine->graph_stats(g);
        out4[0] = c->engine->ticks;
        out4[1] = g[0];
        out4[2] = g[1];
        out4[3] = c->engine->peak_inflight;
        return PFEM_OK;
    });
}

int pfem_case_setup_times(pfem_case* c, double* out6) {
    return guarded([&] {
        PFEM_REQUIRE(c != nullptr && out6 != nullptr, PFEM_E_INVALID, "null argument");
        for (int k = 0; k < 4; ++k) out6[k] = c->setup_s[k];
        out6[4] = c->min_quality;
        out6[5] = std::min(c->amg_omega_cap, 0.9);
        return PFEM_OK;
    });
}

int pfem_case_workspace_bytes(pfem_case* c, int64_t* bytes) {
    return guarded([&] {
        PFEM_REQUIRE(c != nullptr && bytes != nullptr, PFEM_E_INVALID, "null argument");
        *bytes = (int64_t)pfem::workspace_bytes(c);
        return PFEM_OK;
    });
}

int pfem_case_reserve_workspace(pfem_case* c) {
    return guarded([&] {
        require_mesh(c);
        PFEM_HIP(hipSetDevice(c->ctx->device));
        pfem::own_workspace(c, c->stream);
        case_engine(c);
        PFEM_HIP(hipStreamSynchronize(c->stream));
        return PFEM_OK;
    });
}

int pfem_case_release_workspace(pfem_case* c) {
    return guarded([&] {
        require_case(c);
        PFEM_HIP(hipStreamSynchronize(c->stream));
        c->ws.clear();
        return PFEM_OK;
    });
}

int pfem_pspg_assemble(pfem_case* c, double dt, const double* q_prev, const double* q_iter) {
    return guarded([&] {
        require_pspg(c);
        PFEM_REQUIRE(dt > 0, PFEM_E_INVALID, "dt must be > 0");
        PFEM_HIP(hipSetDevice(c->ctx->device));
        pfem::own_workspace(c, c->stream);
        const size_t nb = sizeof(double) * (size_t)c->n;
        if (q_prev) PFEM_HIP(hipMemcpyAsync(c->q_prev.p, q_prev, nb, hipMemcpyHostToDevice, c->stream));
        if (q_iter) PFEM_HIP(hipMemcpyAsync(c->ws.x, q_iter, nb, hipMemcpyHostToDevice, c->stream));
        Req r;
        r.kind = pfem::RK_ASSEMBLE;
        r.dt = dt;
        return run_direct(c, r);
    });
}

int pfem_pspg_solve(pfem_case* c, const double* x0, double* x_out, pfem_solve_stats* stats) {
    return guarded([&] {
        require_system(c);
        PFEM_HIP(hipSetDevice(c->ctx->device));
        const size_t nb = sizeof(double) * (size_t)c->n;
        if (x0) PFEM_HIP(hipMemcpyAsync(c->ws.x, x0, nb, hipMemcpyHostToDevice, c->stream));
        Req r;
        r.kind = pfem::RK_SOLVE;
        r.dt = 1.0;
        const int rc = run_direct(c, r);
        if (x_out) PFEM_HIP(hipMemcpyAsync(x_out, c->ws.x, nb, hipMemcpyDeviceToHost, c->stream));
        PFEM_HIP(hipStreamSynchronize(c->stream));
        if (stats) *stats = r.ss;
        return rc;
    });
}

int pfem_get_csr_sizes(pfem_case* c, int64_t* n, int64_t* nnz) {
    return guarded([&] {
        require_pspg(c);
        if (n) *n = c->n;
        if (nnz) *nnz = c->pat.nnz;
        return PFEM_OK;
    });
}

int pfem_get_csr(pfem_case* c, int32_t* row_ptr, int32_t* col_idx, double* vals) {
    return guarded([&] {
        require_pspg(c);
        PFEM_HIP(hipSetDevice(c->ctx->device));
        if (row_ptr)
            PFEM_HIP(hipMemcpyAsync(row_ptr, c->pat.row_ptr.p, sizeof(int) * (size_t)(c->n + 1), hipMemcpyDeviceToHost,
                                    c->stream));
        DevBuf cols;
        if (col_idx) {
            cols.reserve(sizeof(int) * (size_t)c->pat.nnz);
            pfem::build_col_idx(c->pat, cols.as<int>(), c->stream);
            PFEM_HIP(hipMemcpyAsync(col_idx, cols.p, sizeof(int) * (size_t)c->pat.nnz, hipMemcpyDeviceToHost,
                                    c->stream));
        }
        if (vals) {
            PFEM_REQUIRE(c->ws.base != nullptr, PFEM_E_STATE, "no assembled system on this case (assemble first)");
            PFEM_HIP(hipMemcpyAsync(vals, c->ws.vals, sizeof(double) * (size_t)c->pat.nnz, hipMemcpyDeviceToHost,
                                    c->stream));
        }
        PFEM_HIP(hipStreamSynchronize(c->stream));
        return PFEM_OK;
    });
}

int pfem_get_rhs(pfem_case* c, double* b) {
    return guarded([&] {
        require_system(c);
        PFEM_REQUIRE(b != nullptr, PFEM_E_INVALID, "null output");
        PFEM_HIP(hipMemcpyAsync(b, c->ws.b, sizeof(double) * (size_t)c->n, hipMemcpyDeviceToHost, c->stream));
        PFEM_HIP(hipStreamSynchronize(c->stream));
        return PFEM_OK;
    });
}

int pfem_spmv(pfem_case* c, const double* x, double* y) {
    return guarded([&] {
        require_system(c);
        PFEM_REQUIRE(x && y, PFEM_E_INVALID, "null vector");
        PFEM_HIP(hipSetDevice(c->ctx->device));
        const size_t nb = sizeof(double) * (size_t)c->n;
        double* dx = c->ws.kv64;
        double* dy = dx + c->n;
        PFEM_HIP(hipMemcpyAsync(dx, x, nb, hipMemcpyHostToDevice, c->stream));
        pfem::launch_spmv<double>(c->lanes, node_mat<double>(c->pat, c->dim, c->ws.vals), dx, dy,
                                  grid_for_rows(c->N), c->stream);
        PFEM_HIP(hipGetLastError());
        PFEM_HIP(hipMemcpyAsync(y, dy, nb, hipMemcpyDeviceToHost, c->stream));
        PFEM_HIP(hipStreamSynchronize(c->stream));
        return PFEM_OK;
    });
}

int pfem_spmv_ell(pfem_case* c, const double* x, double* y) {
    return guarded([&] {
        require_system(c);
        PFEM_REQUIRE(x && y, PFEM_E_INVALID, "null vector");
        PFEM_HIP(hipSetDevice(c->ctx->device));
        const size_t nb = sizeof(double) * (size_t)c->n;
        double* dx = c->ws.kv64;
        double* dy = dx + c->n;
        PFEM_HIP(hipMemcpyAsync(dx, x, nb, hipMemcpyHostToDevice, c->stream));
        pfem::EllMatT<double> E{};
        E.N = c->N;
        E.nf = c->dim + 1;
        E.nsl = c->pat.ell_nsl;
        E.soff = c->pat.ell_soff.as<int>();
        E.idx = c->pat.ell_idx.as<int>();
        E.vals = c->ws.vals_ell;
        pfem::launch_spmv_ell(E, dx, dy, std::max(1, std::min(8192, (c->pat.ell_nsl + pfem::kWaves - 1) / pfem::kWaves)),
                              c->stream);
        PFEM_HIP(hipGetLastError());
        PFEM_HIP(hipMemcpyAsync(y, dy, nb, hipMemcpyDeviceToHost, c->stream));
        PFEM_HIP(hipStreamSynchronize(c->stream));
        return PFEM_OK;
    });
}

int pfem_pspg_step(pfem_case* c, double dt, pfem_step_stats* stats) {
    return guarded([&] {
        require_pspg(c);
        require_static_dirichlet(c);
        PFEM_REQUIRE(dt > 0, PFEM_E_INVALID, "dt must be > 0");
        PFEM_HIP(hipSetDevice(c->ctx->device));
        Req r;
        r.kind = pfem::RK_STEP;
        r.dt = dt;
        const int rc = run_direct(c, r);
        if (stats) {
            *stats = r.st;
            stats->status = rc;
        }
        return rc;
    });
}

// ---- the Picard loop's callbacks one by one ----------------------------------------------
namespace {
int run_callback(pfem_case* c, unsigned ops, unsigned copy, double dt, int res_kind, double* res_out) {
    Req r;
    r.kind = pfem::RK_CALLBACK;
    r.cb_ops = ops;
    r.cb_copy = copy;
    r.dt = dt;
    r.res_kind = res_kind;
    const int rc = run_direct(c, r);
    if (res_out) *res_out = r.res;
    return rc;
}
}  // namespace

int pfem_save_positions(pfem_case* c) {
    return guarded([&] {
        require_pspg(c);
        PFEM_HIP(hipSetDevice(c->ctx->device));
        return run_callback(c, pfem::OP_COPY, pfem::COPY_PREP, 1.0, -1, nullptr);
    });
}

int pfem_restore_positions(pfem_case* c) {
    return guarded([&] {
        require_pspg(c);
        PFEM_HIP(hipSetDevice(c->ctx->device));
        return run_callback(c, pfem::OP_COPY, pfem::COPY_RESTORE | pfem::COPY_RESET_X, 1.0, -1, nullptr);
    });
}

int pfem_update_positions(pfem_case* c, double dt, const double* q) {
    return guarded([&] {
        require_pspg(c);
        PFEM_REQUIRE(dt > 0, PFEM_E_INVALID, "dt must be > 0");
        PFEM_HIP(hipSetDevice(c->ctx->device));
        pfem::own_workspace(c, c->stream);
        if (q)
            PFEM_HIP(hipMemcpyAsync(c->ws.x, q, sizeof(double) * (size_t)c->n, hipMemcpyHostToDevice, c->stream));
        return run_callback(c, pfem::OP_UPD, 0, dt, -1, nullptr);
    });
}

int pfem_residual(pfem_case* c, int32_t kind, const double* q_iter, const double* q_iter_prev, double* out) {
    return guarded([&] {
        require_pspg(c);
        PFEM_REQUIRE(out != nullptr, PFEM_E_INVALID, "null output");
        PFEM_REQUIRE(kind == PFEM_RES_U || kind == PFEM_RES_U_P || kind == PFEM_RES_AX_F, PFEM_E_INVALID,
                     "unknown residual kind");
        if (kind == PFEM_RES_AX_F) require_system(c);
        PFEM_HIP(hipSetDevice(c->ctx->device));
        pfem::own_workspace(c, c->stream);
        const size_t nb = sizeof(double) * (size_t)c->n;
        if (q_iter) PFEM_HIP(hipMemcpyAsync(c->ws.x, q_iter, nb, hipMemcpyHostToDevice, c->stream));
        if (q_iter_prev) PFEM_HIP(hipMemcpyAsync(c->ws.x_prev, q_iter_prev, nb, hipMemcpyHostToDevice, c->stream));
        double res = 0.0;
        const int rc = run_callback(c, pfem::OP_RES, 0, 1.0, kind, &res);
        *out = res;
        return rc;
    });
}

int pfem_pspg_commit(pfem_case* c) {
    return guarded([&] {
        require_pspg(c);
        PFEM_HIP(hipSetDevice(c->ctx->device));
        return run_callback(c, pfem::OP_COPY, pfem::COPY_COMMIT, 1.0, -1, nullptr);
    });
}

int pfem_get_iterate(pfem_case* c, double* q) {
    return guarded([&] {
        require_pspg(c);
        PFEM_REQUIRE(q != nullptr, PFEM_E_INVALID, "null output");
        PFEM_HIP(hipSetDevice(c->ctx->device));
        pfem::own_workspace(c, c->stream);
        PFEM_HIP(hipMemcpyAsync(q, c->ws.x, sizeof(double) * (size_t)c->n, hipMemcpyDeviceToHost, c->stream));
        PFEM_HIP(hipStreamSynchronize(c->stream));
        return PFEM_OK;
    });
}

int pfem_wc_step(pfem_case* c, double dt, pfem_step_stats* stats) {
    return guarded([&] {
        require_mesh(c);
        PFEM_REQUIRE(c->d.kind == PFEM_KIND_WC, PFEM_E_STATE, "not a WC case");
        PFEM_REQUIRE(dt > 0, PFEM_E_INVALID, "dt must be > 0");
        PFEM_HIP(hipSetDevice(c->ctx->device));
        Req r;
        r.kind = pfem::RK_WC;
        r.dt = dt;
        r.wc_step = true;
        r.wc_cfl = false;
        r.sc = c->cfl_sc;  // the CFL minimum after the step is kept for this securityCoeff
        const int rc = run_direct(c, r);
        if (stats) {
            *stats = pfem_step_stats{};
            stats->status = rc;
            stats->assemblies = 2;
        }
        return rc;
    });
}

int pfem_wc_next_dt(pfem_case* c, double security_coeff, double max_dt, double* dt_out) {
    return guarded([&] {
        require_mesh(c);
        PFEM_REQUIRE(c->d.kind == PFEM_KIND_WC && dt_out, PFEM_E_INVALID, "not a WC case or null output");
        PFEM_REQUIRE(security_coeff > 0, PFEM_E_INVALID, "securityCoeff must be > 0");
        if (c->cfl_valid && c->cfl_sc == security_coeff) {
            // the last step already reduced the CFL minimum of this state
            const double dt = std::min(std::sqrt(c->cfl_dt2), max_dt);
            PFEM_REQUIRE(!std::isnan(dt), PFEM_NAN, "NaN time step");
            *dt_out = dt;
            return PFEM_OK;
        }
        PFEM_HIP(hipSetDevice(c->ctx->device));
        Req r;
        r.kind = pfem::RK_WC;
        r.dt = 1.0;
        r.wc_step = false;
        r.wc_cfl = true;
        r.sc = security_coeff;
        r.max_dt = max_dt;
        run_direct(c, r);
        PFEM_REQUIRE(!std::isnan(r.next_dt), PFEM_NAN, "NaN time step");
        *dt_out = r.next_dt;
        return PFEM_OK;
    });
}

int pfem_case_mass(pfem_case* c, double* mass) {
    return guarded([&] {
        require_mesh(c);
        PFEM_REQUIRE(mass != nullptr, PFEM_E_INVALID, "null output");
        PFEM_HIP(hipSetDevice(c->ctx->device));
        const int nb = (int)std::min<long long>(((long long)c->E + pfem::kBlock - 1) / pfem::kBlock, pfem::kMaxCaseBlocks);
        const void* rho = nullptr;
        const bool fp32 = c->d.wc_state == PFEM_STATE_FP32;
        if (c->d.kind == PFEM_KIND_WC)  // state layout v (dim x N), p (N), rho (N), a (dim x N)
            rho = static_cast<const char*>(c->wstate.p) + wc_state_size(c) * (size_t)(c->dim + 1) * c->N;
        pfem::launch_mass(c->dim, c->N, c->E, c->pat.conn.as<int>(), c->pos.as<double>(), rho,
                          c->d.kind == PFEM_KIND_WC && fp32 ? 1 : 0, c->partials.as<double>(), nb, c->stream);
        PFEM_HIP(hipGetLastError());
        std::vector<double> h((size_t)nb);
        PFEM_HIP(hipMemcpyAsync(h.data(), c->partials.p, sizeof(double) * (size_t)nb, hipMemcpyDeviceToHost, c->stream));
        PFEM_HIP(hipStreamSynchronize(c->stream));
        double m = 0.0;
        for (double v : h) m += v;
        *mass = m;
        return PFEM_OK;
    });
}

int pfem_case_mass_noalloc(pfem_case* c, double* d_mass) {
    return guarded([&] {
        require_mesh(c);
        PFEM_REQUIRE(d_mass != nullptr, PFEM_E_INVALID, "null output");
        PFEM_HIP(hipSetDevice(c->ctx->device));
        const int nb = (int)std::min<long long>(((long long)c->E + pfem::kBlock - 1) / pfem::kBlock, pfem::kMaxCaseBlocks);
        PFEM_REQUIRE(nb <= pfem::kScratchPartials, PFEM_E_INVALID,
                     "more workgroups than the scratch holds partials (" + std::to_string(pfem::kScratchPartials) + ")");
        double* part = pfem::scratch_partials(c->ctx->device);
        const void* rho = nullptr;
        const bool fp32 = c->d.wc_state == PFEM_STATE_FP32;
        if (c->d.kind == PFEM_KIND_WC)
            rho = static_cast<const char*>(c->wstate.p) + wc_state_size(c) * (size_t)(c->dim + 1) * c->N;
        pfem::launch_mass(c->dim, c->N, c->E, c->pat.conn.as<int>(), c->pos.as<double>(), rho,
                          c->d.kind == PFEM_KIND_WC && fp32 ? 1 : 0, part, nb, c->stream);
        pfem::launch_sum_partials(part, nb, d_mass, c->stream);
        PFEM_HIP(hipGetLastError());
        return PFEM_OK;
    });
}

int pfem_case_minmax(pfem_case* c, int32_t coordinate, int32_t want_max, double* value) {
    return guarded([&] {
        require_mesh(c);
        PFEM_REQUIRE(value != nullptr && coordinate >= 0 && coordinate < c->dim, PFEM_E_INVALID,
                     "coordinate out of range or null output");
        PFEM_HIP(hipSetDevice(c->ctx->device));
        const int nb = (int)std::min<long long>(((long long)c->N + pfem::kBlock - 1) / pfem::kBlock, pfem::kMaxCaseBlocks);
        pfem::launch_minmax(c->N, c->pat.flags.as<uint8_t>(), c->pos.as<double>() + (size_t)coordinate * c->N,
                            want_max ? 1 : 0, c->partials.as<double>(), nb, c->stream);
        PFEM_HIP(hipGetLastError());
        std::vector<double> h((size_t)nb);
        PFEM_HIP(hipMemcpyAsync(h.data(), c->partials.p, sizeof(double) * (size_t)nb, hipMemcpyDeviceToHost, c->stream));
        PFEM_HIP(hipStreamSynchronize(c->stream));
        double v = want_max ? 0.0 : std::numeric_limits<double>::max();
        for (double x : h) v = want_max ? std::max(v, x) : std::min(v, x);
        *value = v;
        return PFEM_OK;
    });
}

int pfem_profile_spmv(pfem_case* c, int32_t enable, double* out4) {
    return guarded([&] {
        require_case(c);
        pfem::Engine& E = case_engine(c);
        if (out4) {
            for (int k = 0; k < 2; ++k) {
                out4[2 * k] = E.prof_count[k] ? E.prof_ms[k] / (double)E.prof_count[k] : 0.0;
                out4[2 * k + 1] = (double)E.prof_count[k];
            }
        }
        E.enable_profiling(enable != 0);
        return PFEM_OK;
    });
}

int pfem_bench_kernels(pfem_case* c, int32_t reps, double* out10) {
    return guarded([&] {
        require_system(c);
        PFEM_REQUIRE(out10 && reps > 0, PFEM_E_INVALID, "invalid arguments");
        double* out8 = out10;
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
            float ms = 0;
            PFEM_HIP(hipEventElapsedTime(&ms, e0, e1));
            return 1e3 * (double)ms / reps;  // microseconds per launch
        };
        const int n = c->n;
        double* w = c->ws.kv64;
        for (int k = 0; k < 10; ++k) out10[k] = 0;
        {  // the BiCGSTAB SpMV kernel (v = A ph with the fused (rh, v)) as a solve launches it
            const int nb = pfem::row_blocks(n, c->pat.ell_nsl);
            std::vector<pfem::BlockTask> tasks((size_t)nb);
            for (int b = 0; b < nb; ++b) tasks[(size_t)b] = pfem::BlockTask{0, b, nb, 0};
            pfem::KrylovView<double> kv{};
            kv.n = n;
            kv.nodes = c->N;
            kv.nf = c->dim + 1;
            kv.nblk = nb;
            kv.A = node_mat<double>(c->pat, c->dim, c->ws.vals);
            kv.E.N = c->N;
            kv.E.nf = c->dim + 1;
            kv.E.nsl = c->pat.ell_nsl;
            kv.E.soff = c->pat.ell_soff.as<int>();
            kv.E.idx = c->pat.ell_idx.as<int>();
            kv.E.vals = c->ws.vals_ell;
            kv.ph = c->q_prev.as<double>();
            kv.rh = w;
            kv.v = w + n;
            kv.partials = c->partials.as<double>();
            kv.counters = c->counters.as<unsigned>();
            DevBuf d_tasks, d_view, d_sc;
            d_tasks.reserve(sizeof(pfem::BlockTask) * tasks.size());
            d_view.reserve(sizeof(kv));
            d_sc.reserve(sizeof(pfem::KScal));
            PFEM_HIP(hipMemsetAsync(d_sc.p, 0, sizeof(pfem::KScal), c->stream));  // done = 0: the kernel runs
            kv.sc = d_sc.as<pfem::KScal>();
            PFEM_HIP(hipMemcpyAsync(d_tasks.p, tasks.data(), sizeof(pfem::BlockTask) * tasks.size(),
                                    hipMemcpyHostToDevice, c->stream));
            PFEM_HIP(hipMemcpyAsync(d_view.p, &kv, sizeof(kv), hipMemcpyHostToDevice, c->stream));
            out8[0] = timed([&] {
                pfem::launch_kry_phase<double>(1, c->lanes, d_view.as<pfem::KrylovView<double>>(),
                                               d_tasks.as<pfem::BlockTask>(), nb, 0, c->stream, nullptr);
            });
        }
        if (c->d.precision == PFEM_PREC_MIXED_IR) {
            float* u = c->ws.kv32;
            out8[1] = timed([&] {
                pfem::launch_spmv<float>(c->lanes, node_mat<float>(c->pat, c->dim, c->ws.vals32), u, u + n,
                                         grid_for_rows(c->N), c->stream);
            });
        }
        const long long nnz = c->pat.nnz;
        out8[2] = timed([&] { pfem::launch_stream_read(c->ws.vals, nnz, w, 2048, c->stream); });
        const long long ncopy = std::min<long long>(nnz, 4LL * n);
        out8[3] = timed([&] { pfem::launch_copy(c->ws.vals, w + 4LL * n, ncopy, 2048, c->stream); });
        out8[4] = (double)ncopy * 16.0;
        out8[5] = (double)nnz;
        out8[6] = (double)c->pat.graph_nnz;
        out8[7] = (double)c->lanes;
        if (c->amg.levels > 1) {  // the multigrid's level-0 residual t = r - A w over the FP64 ELL copy
            const int nb = (int)std::min<long long>((c->N + pfem::kBlock - 1) / pfem::kBlock, 8192);
            std::vector<pfem::BlockTask> tasks((size_t)nb);
            for (int b = 0; b < nb; ++b) tasks[(size_t)b] = pfem::BlockTask{0, b, nb, 0};
            DevBuf d_tasks, d_view, d_done;
            d_done.reserve(sizeof(int));
            PFEM_HIP(hipMemsetAsync(d_done.p, 0, sizeof(int), c->stream));
            PFEM_HIP(hipMemsetAsync(w + 2LL * n, 0, sizeof(double) * (size_t)n, c->stream));
            pfem::AmgView<double> av{};
            av.A = node_mat<double>(c->pat, c->dim, c->ws.vals);
            av.E0.N = c->N;
            av.E0.nf = c->dim + 1;
            av.E0.nsl = c->pat.ell_nsl;
            av.E0.soff = c->pat.ell_soff.as<int>();
            av.E0.idx = c->pat.ell_idx.as<int>();
            av.E0.vals = c->ws.vals_ell;
            av.level = 0;
            av.lanes = c->lanes;
            av.mask = 1;
            av.flags = c->pat.flags.as<uint8_t>();
            av.done_flag = d_done.as<int>();
            av.r0[0] = c->q_prev.as<double>();
            av.w = w + 2LL * n;
            av.t = w + 3LL * n;
            d_tasks.reserve(sizeof(pfem::BlockTask) * tasks.size());
            d_view.reserve(sizeof(av));
            PFEM_HIP(hipMemcpyAsync(d_tasks.p, tasks.data(), sizeof(pfem::BlockTask) * tasks.size(),
                                    hipMemcpyHostToDevice, c->stream));
            PFEM_HIP(hipMemcpyAsync(d_view.p, &av, sizeof(av), hipMemcpyHostToDevice, c->stream));
            out10[8] = timed([&] {
                pfem::launch_amg_res<double>(d_view.as<pfem::AmgView<double>>(), d_tasks.as<pfem::BlockTask>(), nb, 0,
                                             c->stream);
            });
            out10[9] = (double)nb;
        }
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
        std::vector<Req> reqs((size_t)ncases);
        for (int i = 0; i < ncases; ++i) {
            pfem_case* c = cases[i];
            require_mesh(c);
            PFEM_REQUIRE(c->d.kind == PFEM_KIND_PSPG || c->d.kind == PFEM_KIND_WC, PFEM_E_UNSUPPORTED,
                         "heat, fractional-step and Boussinesq cases are stepped with pfem_heat_step / "
                         "pfem_frac_step / pfem_heat_step + pfem_pspg_step");
            if (c->d.kind == PFEM_KIND_PSPG) require_static_dirichlet(c);
            PFEM_REQUIRE(c->ctx == ctx, PFEM_E_INVALID, "case belongs to another context");
            PFEM_REQUIRE(dts[i] > 0, PFEM_E_INVALID, "dt must be > 0");
            for (int j = 0; j < i; ++j) PFEM_REQUIRE(cases[j] != c, PFEM_E_INVALID, "a case is listed twice");
            PFEM_HIP(hipStreamSynchronize(c->stream));
            Req& r = reqs[i];
            r.c = c;
            r.dt = dts[i];
            r.kind = c->d.kind == PFEM_KIND_PSPG ? pfem::RK_STEP : pfem::RK_WC;
            r.wc_step = true;
            r.wc_cfl = false;
            r.sc = c->cfl_sc;
        }
        if (!ctx->engine) ctx->engine = std::make_unique<pfem::Engine>(ctx->device, ctx->stream, std::max(1, ncases),
                                                                       nullptr);
        pfem::Engine& E = *ctx->engine;
        E.ensure_slots(ncases);
        std::vector<Req*> done;
        const int64_t t0 = pfem::now_ns();
        for (auto& r : reqs) {
            r.t_arrival = t0;
            E.enqueue(&r);
        }
        try {
            E.drain(done);
        } catch (...) {
            std::vector<Req*> d2;
            E.abort_all(PFEM_E_HIP, d2);
            throw;
        }
        for (int i = 0; i < ncases; ++i) {
            if (status) status[i] = reqs[i].status;
            if (stats) {
                stats[i] = reqs[i].st;
                stats[i].status = reqs[i].status;
                if (reqs[i].kind == pfem::RK_WC) stats[i].assemblies = 2;
            }
        }
        return PFEM_OK;
    });
}

// ---- serving ---------------------------------------------------------------------------
int pfem_server_create(pfem_ctx* ctx, const pfem_server_desc* desc, pfem_server** out) {
    return guarded([&] {
        PFEM_REQUIRE(ctx && desc && out, PFEM_E_INVALID, "null argument");
        PFEM_REQUIRE(desc->max_inflight > 0 && desc->max_inflight <= 4096, PFEM_E_INVALID,
                     "max_inflight must be in [1, 4096]");
        PFEM_REQUIRE(desc->hbm_budget >= 0, PFEM_E_INVALID, "negative budget");
        PFEM_HIP(hipSetDevice(ctx->device));
        auto S = std::make_unique<pfem_server>();
        S->ctx = ctx;
        size_t budget = (size_t)desc->hbm_budget;
        if (budget == 0) {
            size_t fr = 0, tot = 0;
            PFEM_HIP(hipMemGetInfo(&fr, &tot));
            budget = fr / 2;
        }
        S->arena.init(budget);
        PFEM_HIP(hipStreamCreateWithFlags(&S->stream, hipStreamNonBlocking));
        S->eng = std::make_unique<pfem::Engine>(ctx->device, S->stream, desc->max_inflight, &S->arena);
        if (desc->max_chunk > 0) S->eng->max_chunk = std::max(4, std::min(64, desc->max_chunk));
        S->wc_sc = desc->wc_security_coeff > 0 ? desc->wc_security_coeff : 0.5;
        S->wc_max_dt = desc->wc_max_dt > 0 ? desc->wc_max_dt : 1.0;
        pfem_server* raw = S.release();
        raw->th = std::thread(server_loop, raw);
        *out = raw;
        return PFEM_OK;
    });
}

int pfem_server_destroy(pfem_server* S) {
    return guarded([&] {
        if (!S) return PFEM_OK;
        {
            std::lock_guard<std::mutex> lk(S->mu);
            S->stop = true;
        }
        S->cv.notify_all();
        if (S->th.joinable()) S->th.join();
        for (Req* r : S->results) delete r;
        S->eng.reset();
        (void)hipStreamDestroy(S->stream);
        delete S;
        return PFEM_OK;
    });
}

int pfem_server_submit(pfem_server* S, pfem_case* c, double dt, int64_t tag, int64_t arrival_ns) {
    return guarded([&] {
        PFEM_REQUIRE(S && c, PFEM_E_INVALID, "null argument");
        PFEM_REQUIRE(c->has_mesh, PFEM_E_STATE, "no mesh set on this case");
        PFEM_REQUIRE(c->ctx == S->ctx, PFEM_E_INVALID, "case belongs to another context");
        PFEM_REQUIRE(dt > 0, PFEM_E_INVALID, "dt must be > 0");
        PFEM_REQUIRE(c->d.kind == PFEM_KIND_PSPG || c->d.kind == PFEM_KIND_WC, PFEM_E_UNSUPPORTED,
                     "heat, fractional-step and Boussinesq cases are stepped with pfem_heat_step / "
                     "pfem_frac_step / pfem_heat_step + pfem_pspg_step");
        if (c->d.kind == PFEM_KIND_PSPG) require_static_dirichlet(c);
        const size_t wb = c->ws.base ? 0 : pfem::workspace_bytes(c);
        PFEM_REQUIRE(wb <= S->arena.capacity(), PFEM_E_OOM, "the case's step workspace exceeds the server budget");
        int expected = 0;
        PFEM_REQUIRE(c->busy.compare_exchange_strong(expected, 1), PFEM_E_STATE, "case already has a pending request");
        PFEM_HIP(hipStreamSynchronize(c->stream));
        auto r = std::make_unique<Req>();
        r->c = c;
        r->dt = dt;
        r->tag = tag;
        if (c->d.kind == PFEM_KIND_PSPG) {
            r->kind = pfem::RK_STEP;
        } else {
            // a WC step followed by computeNextDT, as the reference's time loop does
            r->kind = pfem::RK_WC;
            r->wc_step = true;
            r->wc_cfl = true;
            r->sc = S->wc_sc;
            r->max_dt = S->wc_max_dt;
        }
        r->t_arrival = arrival_ns > 0 ? arrival_ns : pfem::now_ns();
        {
            std::lock_guard<std::mutex> lk(S->mu);
            if (S->err || S->stop) {
                c->busy.store(0);
                PFEM_REQUIRE(false, S->err ? S->err : PFEM_E_STATE, "server stopped: " + S->err_msg);
            }
            S->incoming.emplace(r->t_arrival, r.release());
            S->submitted++;
        }
        S->cv.notify_one();
        return PFEM_OK;
    });
}

int pfem_server_poll(pfem_server* S, pfem_server_result* out, int32_t max, int32_t timeout_ms, int32_t* n_out) {
    return guarded([&] {
        PFEM_REQUIRE(S && n_out && (max == 0 || out), PFEM_E_INVALID, "null argument");
        std::unique_lock<std::mutex> lk(S->mu);
        if (S->results.empty() && timeout_ms != 0) {
            auto pred = [&] { return !S->results.empty() || S->stop; };
            if (timeout_ms < 0) S->cv_done.wait(lk, pred);
            else S->cv_done.wait_for(lk, std::chrono::milliseconds(timeout_ms), pred);
        }
        int k = 0;
        while (k < max && !S->results.empty()) {
            Req* r = S->results.front();
            S->results.pop_front();
            pfem_server_result& o = out[k++];
            o = pfem_server_result{};
            o.tag = r->tag;
            o.status = r->status;
            o.kind = r->c->d.kind;
            o.stats = r->st;
            o.stats.status = r->status;
            if (r->kind == pfem::RK_WC) o.stats.assemblies = 2;
            o.next_dt = r->next_dt;
            o.t_arrival_ns = r->t_arrival;
            o.t_admit_ns = r->t_admit;
            o.t_done_ns = r->t_done;
            delete r;
        }
        *n_out = k;
        if (S->err) {
            g_err = S->err_msg;
            return S->err;
        }
        return PFEM_OK;
    });
}

int pfem_server_stats(pfem_server* S, int64_t* out8) {
    return guarded([&] {
        PFEM_REQUIRE(S && out8, PFEM_E_INVALID, "null argument");
        std::lock_guard<std::mutex> lk(S->mu);
        out8[0] = S->ticks;
        out8[1] = S->peak_inflight;
        out8[2] = (int64_t)S->arena.capacity();
        out8[3] = S->arena_peak;
        out8[4] = S->submitted;
        out8[5] = S->completed;
        out8[6] = S->arena_used;
        out8[7] = (int64_t)S->incoming.size();
        return PFEM_OK;
    });
}

// ---- heat (conduction) -----------------------------------------------------------------------
int pfem_heat_set_facets(pfem_case* c, int32_t F, const int32_t* nodes, const int32_t* out_node,
                         const uint8_t* kinds) {
    return guarded([&] {
        require_heat(c);
        PFEM_REQUIRE(F >= 0 && (F == 0 || (nodes && out_node && kinds)), PFEM_E_INVALID, "invalid facets");
        const int d = c->dim, N = c->N;
        for (int64_t q = 0; q < (int64_t)F * d; ++q)
            PFEM_REQUIRE(nodes[q] >= 0 && nodes[q] < N, PFEM_E_INVALID, "facet node out of range");
        for (int32_t f = 0; f < F; ++f)
            PFEM_REQUIRE(out_node[f] >= 0 && out_node[f] < N, PFEM_E_INVALID, "facet out node out of range");
        auto& H = c->heat;
        // node -> (facet, local) entries in facet order
        std::vector<int> cnt((size_t)N + 1, 0), list((size_t)F * d);
        for (int64_t q = 0; q < (int64_t)F * d; ++q) cnt[nodes[q] + 1]++;
        for (int i = 0; i < N; ++i) cnt[i + 1] += cnt[i];
        std::vector<int> fill(cnt.begin(), cnt.end() - 1);
        for (int64_t q = 0; q < (int64_t)F * d; ++q) list[fill[nodes[q]]++] = (int)q;
        auto up = [&](pfem::DevBuf& b, const void* src, size_t bytes) {
            b.reserve(bytes);
            if (bytes) PFEM_HIP(hipMemcpyAsync(b.p, src, bytes, hipMemcpyHostToDevice, c->stream));
        };
        up(H.facets, nodes, sizeof(int) * (size_t)F * d);
        up(H.fout, out_node, sizeof(int) * (size_t)F);
        up(H.fkind, kinds, (size_t)F);
        up(H.nf_ptr, cnt.data(), sizeof(int) * cnt.size());
        up(H.nf_list, list.data(), sizeof(int) * list.size());
        H.fterm.reserve(sizeof(double) * 3 * (size_t)F * d);
        H.F = F;
        H.has_q = false;
        PFEM_HIP(hipStreamSynchronize(c->stream));
        return PFEM_OK;
    });
}

int pfem_heat_facet_points(pfem_case* c, double* xgp) {
    return guarded([&] {
        require_heat(c);
        PFEM_REQUIRE(xgp != nullptr || c->heat.F == 0, PFEM_E_INVALID, "null output");
        const int d = c->dim, N = c->N, F = c->heat.F;
        std::vector<double> pos((size_t)d * N);
        std::vector<int> fn((size_t)F * d);
        PFEM_HIP(hipMemcpyAsync(pos.data(), c->pos.p, sizeof(double) * pos.size(), hipMemcpyDeviceToHost, c->stream));
        if (F) PFEM_HIP(hipMemcpyAsync(fn.data(), c->heat.facets.p, sizeof(int) * fn.size(), hipMemcpyDeviceToHost,
                                       c->stream));
        PFEM_HIP(hipStreamSynchronize(c->stream));
        // Facet::getPosFromGP: x = sum_i N_i(gp) x_i with the facet rule of the heat equation
        const double s35 = std::sqrt(3.0 / 5.0);
        const double N2[3][2] = {{(1 + s35) / 2, (1 - s35) / 2}, {0.5, 0.5}, {(1 - s35) / 2, (1 + s35) / 2}};
        const double gp3[3][2] = {{1.0 / 6.0, 1.0 / 6.0}, {1.0 / 6.0, 2.0 / 3.0}, {2.0 / 3.0, 1.0 / 6.0}};
        for (int f = 0; f < F; ++f)
            for (int g = 0; g < 3; ++g)
                for (int k = 0; k < d; ++k) {
                    double x = 0;
                    for (int i = 0; i < d; ++i) {
                        const double Ni = d == 2 ? N2[g][i]
                                                 : (i == 0 ? 1 - gp3[g][0] - gp3[g][1] : gp3[g][i - 1]);
                        x += Ni * pos[(size_t)k * N + fn[(size_t)f * d + i]];
                    }
                    xgp[((size_t)f * 3 + g) * d + k] = x;
                }
        return PFEM_OK;
    });
}

int pfem_heat_set_flux(pfem_case* c, const double* q) {
    return guarded([&] {
        require_heat(c);
        auto& H = c->heat;
        if (!q || H.F == 0) {
            H.has_q = false;
            return PFEM_OK;
        }
        const size_t bytes = sizeof(double) * (size_t)H.F * 3 * c->dim;
        H.qgp.reserve(bytes);
        PFEM_HIP(hipMemcpyAsync(H.qgp.p, q, bytes, hipMemcpyHostToDevice, c->stream));
        PFEM_HIP(hipStreamSynchronize(c->stream));
        H.has_q = true;
        return PFEM_OK;
    });
}

int pfem_heat_set_dirichlet(pfem_case* c, const double* t_bc) {
    return guarded([&] {
        require_heat(c);
        PFEM_REQUIRE(t_bc != nullptr, PFEM_E_INVALID, "null Dirichlet values");
        PFEM_HIP(hipMemcpyAsync(c->heat.Tbc.p, t_bc, sizeof(double) * (size_t)c->N, hipMemcpyHostToDevice, c->stream));
        PFEM_HIP(hipStreamSynchronize(c->stream));
        return PFEM_OK;
    });
}

int pfem_heat_assemble(pfem_case* c, double dt) {
    return guarded([&] {
        require_heat(c);
        PFEM_REQUIRE(dt > 0, PFEM_E_INVALID, "dt must be > 0");
        PFEM_HIP(hipSetDevice(c->ctx->device));
        pfem::launch_heat_assemble(heat_view(c, dt), c->stream);
        PFEM_HIP(hipGetLastError());
        PFEM_HIP(hipStreamSynchronize(c->stream));
        return PFEM_OK;
    });
}

int pfem_heat_system_nnz(pfem_case* c, int64_t* nnz) {
    return guarded([&] {
        require_heat(c);
        PFEM_REQUIRE(nnz != nullptr, PFEM_E_INVALID, "null output");
        *nnz = c->heat.nnz;
        return PFEM_OK;
    });
}

int pfem_heat_get_system(pfem_case* c, int32_t* row_ptr, int32_t* col_idx, double* vals, double* b) {
    return guarded([&] {
        require_heat(c);
        auto& H = c->heat;
        const size_t N = (size_t)c->N, nnz = (size_t)H.nnz;
        if (row_ptr) PFEM_HIP(hipMemcpyAsync(row_ptr, H.row_ptr.p, sizeof(int) * (N + 1), hipMemcpyDeviceToHost, c->stream));
        if (col_idx) PFEM_HIP(hipMemcpyAsync(col_idx, H.col.p, sizeof(int) * nnz, hipMemcpyDeviceToHost, c->stream));
        if (vals) PFEM_HIP(hipMemcpyAsync(vals, H.vals.p, sizeof(double) * nnz, hipMemcpyDeviceToHost, c->stream));
        if (b) PFEM_HIP(hipMemcpyAsync(b, H.b.p, sizeof(double) * N, hipMemcpyDeviceToHost, c->stream));
        PFEM_HIP(hipStreamSynchronize(c->stream));
        return PFEM_OK;
    });
}

int pfem_heat_step(pfem_case* c, double dt, pfem_step_stats* stats) {
    return guarded([&] {
        require_heat(c);
        PFEM_REQUIRE(dt > 0, PFEM_E_INVALID, "dt must be > 0");
        PFEM_HIP(hipSetDevice(c->ctx->device));
        using clk = std::chrono::steady_clock;
        const auto t0 = clk::now();
        const pfem::HeatView V = heat_view(c, dt);
        pfem::launch_heat_assemble(V, c->stream);
        const auto t1 = clk::now();
        // solveWithGuess(b, qPrev) (HeatEquation.inl:131)
        PFEM_HIP(hipMemcpyAsync(V.x, V.Tprev, sizeof(double) * (size_t)V.N, hipMemcpyDeviceToDevice, c->stream));
        pfem::launch_cg_init(V, c->stream);
        double sc[2] = {0, 0};
        PFEM_HIP(hipMemcpyAsync(sc, V.scal, sizeof sc, hipMemcpyDeviceToHost, c->stream));
        PFEM_HIP(hipStreamSynchronize(c->stream));
        const double bb = sc[1], thr = c->d.krylov_rtol * c->d.krylov_rtol * bb;
        double rr = sc[0];
        int it = 0;
        // iterations in chunks of 8 between host checks of r.r (a chunk may run a few
        // iterations past convergence; CG keeps reducing the residual meanwhile)
        while (bb > 0 && std::isfinite(rr) && rr > thr && it < c->d.krylov_max_iter) {
            const int chunk = std::min(8, c->d.krylov_max_iter - it);
            for (int k = 0; k < chunk; ++k, ++it) pfem::launch_cg_iteration(V, it & 1, c->stream);
            PFEM_HIP(hipMemcpyAsync(sc, V.scal, sizeof sc, hipMemcpyDeviceToHost, c->stream));
            PFEM_HIP(hipStreamSynchronize(c->stream));
            rr = sc[0];
        }
        PFEM_HIP(hipGetLastError());
        int status = PFEM_OK;
        if (!std::isfinite(rr)) status = PFEM_NAN;
        else if (bb > 0 && rr > thr) status = PFEM_NOT_CONVERGED;
        if (status == PFEM_OK) {
            if (bb > 0)
                PFEM_HIP(hipMemcpyAsync(c->heat.T.p, V.x, sizeof(double) * (size_t)V.N, hipMemcpyDeviceToDevice, c->stream));
            else
                PFEM_HIP(hipMemsetAsync(c->heat.T.p, 0, sizeof(double) * (size_t)V.N, c->stream));
        }
        PFEM_HIP(hipStreamSynchronize(c->stream));
        const auto t2 = clk::now();
        if (stats) {
            *stats = pfem_step_stats{};
            stats->status = status;
            stats->picard_iters = 1;
            stats->krylov_iters = it;
            stats->assemblies = 1;
            stats->max_relres = bb > 0 ? std::sqrt(rr / bb) : 0.0;
            stats->residual = stats->max_relres;
            stats->t_assemble = std::chrono::duration<double>(t1 - t0).count();
            stats->t_solve = std::chrono::duration<double>(t2 - t1).count();
        }
        return status;
    });
}

// ---- surface-tension facets ---------------------------------------------------------------------
int pfem_case_set_facets(pfem_case* c, int32_t F, const int32_t* nodes, const int32_t* elem) {
    return guarded([&] {
        require_mesh(c);
        PFEM_REQUIRE(pfem::pspg_like(c) || c->d.kind == PFEM_KIND_WC, PFEM_E_STATE,
                     "surface tension applies to PSPG, Boussinesq and WC cases");
        PFEM_REQUIRE(F >= 0 && (F == 0 || (nodes && elem)), PFEM_E_INVALID, "invalid facets");
        const int D = c->dim, ne = D + 1, N = c->N, E = c->E;
        c->fst.F = 0;
        if (F == 0) return PFEM_OK;
        PFEM_HIP(hipSetDevice(c->ctx->device));
        PFEM_REQUIRE((long long)F < (1LL << 30), PFEM_E_INVALID, "too many facets");
        // the element connectivity (AoS) to check that each facet lies on its element
        std::vector<int> soa((size_t)ne * E);
        PFEM_HIP(hipMemcpyAsync(soa.data(), c->pat.conn.p, sizeof(int) * soa.size(), hipMemcpyDeviceToHost, c->stream));
        PFEM_HIP(hipStreamSynchronize(c->stream));
        std::vector<int> cnt((size_t)N + 1, 0);
        for (int32_t f = 0; f < F; ++f) {
            const int e = elem[f];
            PFEM_REQUIRE(e >= 0 && e < E, PFEM_E_INVALID, "facet element out of range");
            for (int k = 0; k < D; ++k) {
                const int n = nodes[(size_t)f * D + k];
                bool in = false;
                for (int a = 0; a < ne; ++a) in = in || soa[(size_t)a * E + e] == n;
                PFEM_REQUIRE(n >= 0 && n < N && in, PFEM_E_INVALID, "a facet node is not a node of its element");
            }
            for (int a = 0; a < ne; ++a) cnt[(size_t)soa[(size_t)a * E + e] + 1]++;
        }
        for (int i = 0; i < N; ++i) cnt[i + 1] += cnt[i];
        std::vector<int> cur(cnt.begin(), cnt.end() - 1);
        std::vector<unsigned> ent((size_t)cnt[N]);
        for (int32_t f = 0; f < F; ++f)  // facet order within each node's list
            for (int a = 0; a < ne; ++a) {
                const int n = soa[(size_t)a * E + elem[f]];
                ent[(size_t)cur[n]++] = ((unsigned)f << 2) | (unsigned)a;
            }
        auto up = [&](pfem::DevBuf& d, const void* src, size_t bytes) {
            d.reserve(bytes);
            PFEM_HIP(hipMemcpyAsync(d.p, src, bytes, hipMemcpyHostToDevice, c->stream));
        };
        up(c->fst.nodes, nodes, sizeof(int) * (size_t)F * D);
        up(c->fst.elem, elem, sizeof(int) * (size_t)F);
        up(c->fst.ptr, cnt.data(), sizeof(int) * cnt.size());
        up(c->fst.ent, ent.data(), sizeof(unsigned) * std::max<size_t>(ent.size(), 1));
        PFEM_HIP(hipStreamSynchronize(c->stream));
        c->fst.F = F;
        return PFEM_OK;
    });
}

// ---- fractional step ---------------------------------------------------------------------------
int pfem_frac_step(pfem_case* c, double dt, pfem_step_stats* stats) {
    return guarded([&] {
        require_frac(c);
        require_static_dirichlet(c);
        PFEM_REQUIRE(dt > 0, PFEM_E_INVALID, "dt must be > 0");
        PFEM_HIP(hipSetDevice(c->ctx->device));
        using clk = std::chrono::steady_clock;
        auto secs = [](clk::time_point a, clk::time_point b) { return std::chrono::duration<double>(b - a).count(); };
        const int D = c->dim, N = c->N;
        const size_t vN = sizeof(double) * (size_t)D * N, pN = sizeof(double) * (size_t)N;
        auto& F = c->frac;
        const pfem::FracView V = frac_view(c, dt);
        hipStream_t s = c->stream;
        // saveNodesList (:478) and the first iterate's "previous" = the step's start state
        PFEM_HIP(hipMemcpyAsync(F.pos_save.p, c->pos.p, vN, hipMemcpyDeviceToDevice, s));
        PFEM_HIP(hipMemcpyAsync(F.v_it.p, c->q_prev.p, vN, hipMemcpyDeviceToDevice, s));
        PFEM_HIP(hipMemcpyAsync(F.v_itp.p, c->q_prev.p, vN, hipMemcpyDeviceToDevice, s));
        PFEM_HIP(hipMemcpyAsync(F.p_itp.p, V.p_prev, pN, hipMemcpyDeviceToDevice, s));
        const pfem::CgView cg_mk = frac_cg(c, 2, V.mk, V.bv, V.dinv_mk, V.vt);
        const pfem::CgView cg_l = frac_cg(c, 0, V.lv, V.bp, V.dinv_l, V.pn);
        const pfem::CgView cg_m = frac_cg(c, 1, V.ms, V.bv, V.dinv_m, V.dv);
        pfem_step_stats st{};
        int status = PFEM_OK, it = 0;
        double res = std::numeric_limits<double>::max();
        std::vector<double> part((size_t)4 * V.nb);
        auto restore = [&]() {
            PFEM_HIP(hipMemcpyAsync(c->pos.p, F.pos_save.p, vN, hipMemcpyDeviceToDevice, s));
        };
        while (res > c->d.min_res) {
            if (it > c->d.max_iter) {
                restore();
                status = PFEM_NOT_CONVERGED;
                break;
            }
            auto t0 = clk::now();
            pfem::launch_frac_assemble(V, s);
            PFEM_HIP(hipGetLastError());
            st.assemblies++;
            auto t1 = clk::now();
            st.t_assemble += secs(t0, t1);
            int ci = 0;
            double rel = 0;
            int rc = frac_cg_solve(c, cg_mk, &ci, &rel);
            st.krylov_iters += ci;
            st.max_relres = std::max(st.max_relres, rel);
            if (rc == PFEM_OK) {
                auto t2 = clk::now();
                st.t_solve += secs(t1, t2);
                pfem::launch_frac_pcorr(V, s);
                auto t3 = clk::now();
                st.t_assemble += secs(t2, t3);
                rc = frac_cg_solve(c, cg_l, &ci, &rel);
                st.krylov_iters += ci;
                st.max_relres = std::max(st.max_relres, rel);
                t1 = clk::now();
                st.t_solve += secs(t3, t1);
            }
            if (rc == PFEM_OK) {
                pfem::launch_frac_vstep(V, s);
                auto t2 = clk::now();
                st.t_assemble += secs(t1, t2);
                rc = frac_cg_solve(c, cg_m, &ci, &rel);
                st.krylov_iters += ci;
                st.max_relres = std::max(st.max_relres, rel);
                st.t_solve += secs(t2, clk::now());
            }
            if (rc != PFEM_OK) {
                restore();
                status = rc;
                break;
            }
            auto tu = clk::now();
            pfem::launch_frac_update(V, s);
            PFEM_HIP(hipGetLastError());
            PFEM_HIP(hipMemcpyAsync(part.data(), V.res_part, sizeof(double) * part.size(), hipMemcpyDeviceToHost, s));
            PFEM_HIP(hipStreamSynchronize(s));
            auto tr = clk::now();
            st.t_update += secs(tu, tr);
            double acc[4] = {0, 0, 0, 0};
            for (int k = 0; k < 4; ++k)
                for (int b = 0; b < V.nb; ++b) acc[k] += part[(size_t)k * V.nb + b];
            const double big = std::numeric_limits<double>::max();
            const double resV = acc[1] == 0 ? big : std::sqrt(acc[0] / acc[1]);
            double resP = 0;
            if (c->d.residual == PFEM_RES_U_P) resP = acc[3] == 0 ? big : std::sqrt(acc[2] / acc[3]);
            res = std::max(resV, resP);
            if (std::isnan(acc[0]) || std::isnan(acc[1]) || std::isnan(acc[2]) || std::isnan(acc[3])) res = NAN;
            st.t_residual += secs(tr, clk::now());
            if (std::isnan(res)) {
                restore();
                status = PFEM_NAN;
                it++;
                break;
            }
            it++;
        }
        if (status == PFEM_OK) {
            PFEM_HIP(hipMemcpyAsync(c->q_prev.p, F.v_it.p, vN, hipMemcpyDeviceToDevice, s));
            PFEM_HIP(hipMemcpyAsync(c->q_prev.as<double>() + (size_t)D * N, F.pn.p, pN, hipMemcpyDeviceToDevice, s));
        }
        PFEM_HIP(hipStreamSynchronize(s));
        if (stats) {
            st.status = status;
            st.picard_iters = it;
            st.residual = res;
            *stats = st;
        }
        return status;
    });
}

// ---- host remesher ----------------------------------------------------------------------------
struct pfem_mesher {
    pfem::mesh::Mesher m;
    explicit pfem_mesher(const pfem::mesh::MesherParams& p) : m(p) {}
};

int pfem_mesher_create(const pfem_mesher_desc* d, pfem_mesher** out) {
    return guarded([&] {
        PFEM_REQUIRE(d && out, PFEM_E_INVALID, "null argument");
        PFEM_REQUIRE(d->dim == 2 || d->dim == 3, PFEM_E_INVALID, "dim must be 2 or 3");
        PFEM_REQUIRE(d->hchar > 0 && d->alpha > 0 && d->omega > 0 && d->gamma >= 0, PFEM_E_INVALID,
                     "hchar, alpha, omega must be > 0 and gamma >= 0");
        PFEM_REQUIRE(d->n_exclusion_zones >= 0 && (d->n_exclusion_zones == 0 || d->exclusion_zones), PFEM_E_INVALID,
                     "invalid exclusion zones");
        pfem::mesh::MesherParams p;
        p.dim = d->dim;
        p.hchar = d->hchar;
        p.alpha = d->alpha;
        p.omega = d->omega;
        p.gamma = d->gamma;
        p.add_on_fs = d->add_on_fs != 0;
        p.delete_flying_nodes = d->delete_flying_nodes != 0;
        p.laplacian_smoothing_boundaries = d->laplacian_smoothing_boundaries != 0;
        p.has_bounding_box = d->has_bounding_box != 0;
        for (int k = 0; k < 2 * d->dim; ++k) p.bounding_box[k] = d->bounding_box[k];
        for (int z = 0; z < d->n_exclusion_zones; ++z) {
            std::array<double, 6> zone{};
            for (int k = 0; k < 2 * d->dim; ++k) zone[k] = d->exclusion_zones[(size_t)z * 2 * d->dim + k];
            p.exclusion_zones.push_back(zone);
        }
        *out = new pfem_mesher(p);
        return PFEM_OK;
    });
}

int pfem_mesher_destroy(pfem_mesher* m) {
    delete m;
    return PFEM_OK;
}

int pfem_mesher_set_nodes(pfem_mesher* m, int32_t n, int32_t nstates, const double* pos, const double* states,
                          const uint8_t* flags, const int32_t* tags) {
    return guarded([&] {
        PFEM_REQUIRE(m && pos && flags, PFEM_E_INVALID, "null argument");
        PFEM_REQUIRE(n > m->m.params().dim && nstates >= 0, PFEM_E_INVALID, "invalid node count");
        m->m.set_nodes(n, nstates, pos, states, flags, tags);
        return PFEM_OK;
    });
}

int pfem_mesher_set_positions_states(pfem_mesher* m, const double* pos, const double* states) {
    return guarded([&] {
        PFEM_REQUIRE(m != nullptr, PFEM_E_INVALID, "null mesher");
        m->m.set_positions_states(pos, states);
        return PFEM_OK;
    });
}

int pfem_mesher_triangulate(pfem_mesher* m) {
    return guarded([&] {
        PFEM_REQUIRE(m != nullptr, PFEM_E_INVALID, "null mesher");
        PFEM_REQUIRE(m->m.n_nodes() > 0, PFEM_E_STATE, "no nodes set");
        m->m.triangulate();
        return PFEM_OK;
    });
}

int pfem_mesher_remesh(pfem_mesher* m) {
    return guarded([&] {
        PFEM_REQUIRE(m != nullptr, PFEM_E_INVALID, "null mesher");
        PFEM_REQUIRE(m->m.n_elems() > 0, PFEM_E_STATE, "triangulate the mesh before remeshing it");
        m->m.remesh();
        return PFEM_OK;
    });
}

int pfem_mesher_sizes(pfem_mesher* m, int32_t* n_nodes, int32_t* n_elems, int32_t* n_facets) {
    return guarded([&] {
        PFEM_REQUIRE(m != nullptr, PFEM_E_INVALID, "null mesher");
        if (n_nodes) *n_nodes = m->m.n_nodes();
        if (n_elems) *n_elems = m->m.n_elems();
        if (n_facets) *n_facets = m->m.n_facets();
        return PFEM_OK;
    });
}

int pfem_mesher_get_nodes(pfem_mesher* m, double* pos, double* states, uint8_t* flags, int32_t* tags) {
    return guarded([&] {
        PFEM_REQUIRE(m != nullptr, PFEM_E_INVALID, "null mesher");
        m->m.get_nodes(pos, states, flags, tags);
        return PFEM_OK;
    });
}

int pfem_mesher_get_elements(pfem_mesher* m, int32_t* conn) {
    return guarded([&] {
        PFEM_REQUIRE(m && conn, PFEM_E_INVALID, "null argument");
        std::copy(m->m.elems().begin(), m->m.elems().end(), conn);
        return PFEM_OK;
    });
}

int pfem_mesher_get_facets(pfem_mesher* m, int32_t* nodes, int32_t* out_node, int32_t* elem) {
    return guarded([&] {
        PFEM_REQUIRE(m != nullptr, PFEM_E_INVALID, "null mesher");
        if (nodes) std::copy(m->m.facets().begin(), m->m.facets().end(), nodes);
        if (out_node) std::copy(m->m.facet_out().begin(), m->m.facet_out().end(), out_node);
        if (elem) std::copy(m->m.facet_elem().begin(), m->m.facet_elem().end(), elem);
        return PFEM_OK;
    });
}

int pfem_mesher_stats(pfem_mesher* m, double* out8) {
    return guarded([&] {
        PFEM_REQUIRE(m && out8, PFEM_E_INVALID, "null argument");
        const auto& M = m->m;
        out8[0] = M.added_;
        out8[1] = M.removed_;
        out8[2] = M.out_of_box_;
        out8[3] = M.flying_deleted_;
        out8[4] = M.skipped_;
        out8[5] = M.t_nodes_;
        out8[6] = M.t_delaunay_;
        out8[7] = M.t_alpha_;
        return PFEM_OK;
    });
}

static int mesher_states_per_node(pfem_case* c) {
    if (c->d.kind == PFEM_KIND_PSPG || c->d.kind == PFEM_KIND_FRACSTEP) return c->dim + 1;
    if (c->d.kind == PFEM_KIND_BOUSSINESQ) return c->dim + 2;
    return c->d.kind == PFEM_KIND_HEAT ? 1 : 2 * c->dim + 2;
}

int pfem_case_set_mesh_from(pfem_case* c, pfem_mesher* m) {
    return guarded([&] {
        require_case(c);
        PFEM_REQUIRE(m != nullptr, PFEM_E_INVALID, "null mesher");
        const auto& M = m->m;
        PFEM_REQUIRE(M.params().dim == c->dim, PFEM_E_INVALID, "mesher and case dimensions differ");
        PFEM_REQUIRE(M.n_states() == mesher_states_per_node(c), PFEM_E_INVALID,
                     "the mesher's node states do not match the case kind");
        PFEM_REQUIRE(M.n_elems() > 0, PFEM_E_STATE, "the mesher holds no triangulation");
        const int N = M.n_nodes(), E = M.n_elems(), ns = M.n_states();
        std::vector<double> pos((size_t)c->dim * N), st((size_t)ns * N);
        std::vector<uint8_t> fl(N);
        M.get_nodes(pos.data(), st.data(), fl.data(), nullptr);
        int rc = pfem_case_set_mesh(c, N, E, M.elems().data(), fl.data(), pos.data());
        if (rc != PFEM_OK) return rc;
        rc = pfem_case_set_state(c, st.data());
        if (rc != PFEM_OK) return rc;
        // the mesh's facets for the surface-tension term (Mesh::m_facets after the remesh)
        const bool bq = c->d.kind == PFEM_KIND_BOUSSINESQ;
        if ((pfem::pspg_like(c) || c->d.kind == PFEM_KIND_WC) &&
            !(c->d.gamma < 1e-15 && (!bq || c->d.dgamma_dt < 1e-15))) {
            const int F = M.n_facets();
            if (F > 0) rc = pfem_case_set_facets(c, F, M.facets().data(), M.facet_elem().data());
        }
        return rc;
    });
}

int pfem_case_remesh(pfem_case* c, pfem_mesher* m) {
    return guarded([&] {
        require_mesh(c);
        PFEM_REQUIRE(m != nullptr, PFEM_E_INVALID, "null mesher");
        auto& M = m->m;
        PFEM_REQUIRE(M.params().dim == c->dim, PFEM_E_INVALID, "mesher and case dimensions differ");
        PFEM_REQUIRE(M.n_nodes() == c->N && M.n_states() == mesher_states_per_node(c), PFEM_E_STATE,
                     "the mesher does not hold this case's nodes (set them and triangulate first)");
        PFEM_REQUIRE(M.n_elems() > 0, PFEM_E_STATE, "the mesher holds no triangulation");
        const int N = c->N, ns = M.n_states();
        std::vector<double> pos((size_t)c->dim * N), st((size_t)ns * N);
        int rc = pfem_case_get_positions(c, pos.data());
        if (rc != PFEM_OK) return rc;
        rc = pfem_case_get_state(c, st.data());
        if (rc != PFEM_OK) return rc;
        M.set_positions_states(pos.data(), st.data());
        M.remesh();
        rc = pfem_case_set_mesh_from(c, m);
        return rc;
    });
}

}  // extern "C"
