// Host runtime of the PFEM gfx950 library: contexts, cases, step workspaces, the serving
// arena and the tick engine that drives every time step.
//
// Every computation goes through an Engine.  A request (one PSPG Picard step, one linear
// solve, one assembly, or one weakly compressible step) occupies an engine *slot* while it
// runs.  Each tick the engine builds device work tables over all in-flight requests,
// grouped by the operation each request needs next, launches each operation once for
// its whole group, reads every slot's scalars back with one copy, and advances each
// request's state machine.  New requests are admitted at the start of every tick, so a
// request that arrives while others are on the GPU joins the running batch (continuous
// batching).  The direct C-ABI calls run the same engine with one slot.
#pragma once

#include <atomic>
#include <condition_variable>
#include <cstdint>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/pfem_gpu.h"
#include "amg.hpp"
#include "hostutil.hpp"
#include "launch.hpp"
#include "pattern.hpp"

struct pfem_case;

namespace pfem {

class Arena;

// The per-step workspace of a PSPG case: everything rebuilt or overwritten inside one
// time step (matrix values, rhs, block-Jacobi inverse, Picard iterates, saved positions,
// Krylov vectors).  Either owned by the case (direct calls) or borrowed from a serving
// arena for the duration of one request.
struct Workspace {
    char* base = nullptr;
    size_t bytes = 0;
    Arena* arena = nullptr;  // non-null: borrowed at `off`
    size_t off = 0;
    DevBuf own;
    double* vals = nullptr;
    double* b = nullptr;
    double* binv = nullptr;
    double* x = nullptr;
    double* x_prev = nullptr;
    double* pos_save = nullptr;
    double* kv64 = nullptr;  // 8 n: r, rh, p, v, s, t, ph, sh
    double* vals_ell = nullptr;  // sliced ELL copy of the values (EllMat)
    double* erec = nullptr;      // element records of the assembly (E x 16 doubles)
    double* esv = nullptr;       // per-element previous-velocity sums (E x 4 doubles)
    float* vals32_ell = nullptr;  // its FP32 copy (FP32 multigrid hierarchies)
    float* vals32 = nullptr;
    float* binv32 = nullptr;
    float* kv32 = nullptr;  // 10 n: b, x, r, rh, p, v, s, t, ph, sh
    // multigrid (FP64 or FP32 arrays): level-0 smoother vectors, coarse levels' values /
    // inverses / vectors, coarsest dense inverse
    char* amg_w0 = nullptr;
    char* amg_t0 = nullptr;
    char* amg_dense = nullptr;
    char* amg_lv[kAmgMaxLevels][6] = {};  // {vals, binv, r, w, t, z} of level l >= 1
    char* amg_ap[kAmgMaxLevels] = {};     // A P values of level l against level l+1
    // Forgets the arrays (and frees an owned allocation).
    void clear() {
        own.release();
        base = nullptr;
        bytes = 0;
        arena = nullptr;
        off = 0;
        vals = b = binv = x = x_prev = pos_save = kv64 = vals_ell = erec = esv = nullptr;
        vals32 = binv32 = kv32 = vals32_ell = nullptr;
        amg_w0 = amg_t0 = amg_dense = nullptr;
        for (auto& l : amg_lv)
            for (auto& p : l) p = nullptr;
        for (auto& p : amg_ap) p = nullptr;
    }
};

// Bytes of the step workspace of a PSPG case (0 for weakly compressible cases).
size_t workspace_bytes(const pfem_case* c);
// Points the workspace arrays into [base, base + workspace_bytes).
void carve_workspace(pfem_case* c, char* base);
// Gives the case its own workspace (if it has none) with the Picard iterate x = q_prev.
void own_workspace(pfem_case* c, hipStream_t s);

// First-fit allocator over one device allocation.
class Arena {
public:
    void init(size_t bytes);
    bool alloc(size_t bytes, size_t& off);
    void release(size_t off, size_t bytes);
    char* base() const { return static_cast<char*>(mem_.p); }
    size_t capacity() const { return cap_; }
    size_t used() const { return used_; }
    size_t peak() const { return peak_; }

private:
    DevBuf mem_;
    size_t cap_ = 0, used_ = 0, peak_ = 0;
    std::map<size_t, size_t> free_;  // offset -> bytes
};

// Operations a request asks of the next engine tick (Req::ops).
enum : unsigned {
    OP_COPY = 1u << 0,
    OP_UPD = 1u << 1,
    OP_ASM = 1u << 2,
    OP_RES = 1u << 3,
    OP_IRU = 1u << 4,
    OP_IRR = 1u << 5,
    OP_K64_INIT = 1u << 6,
    OP_K64_RUN = 1u << 7,
    OP_K32_INIT = 1u << 8,
    OP_K32_RUN = 1u << 9,
    OP_WC = 1u << 10,
    OP_CFL = 1u << 11,
};

// RK_CALLBACK: one of the Picard loop's device operations on its own (node save / restore,
// position update, residual, commit), as the reference's callbacks call them one by one.
enum ReqKind : int { RK_ASSEMBLE = 0, RK_SOLVE = 1, RK_STEP = 2, RK_WC = 3, RK_CALLBACK = 4 };

struct Req {
    pfem_case* c = nullptr;
    ReqKind kind = RK_STEP;
    double dt = 0;
    int64_t tag = 0;
    bool wc_step = true, wc_cfl = false;
    double sc = 0.5, max_dt = 0;
    int res_kind = -1;                    // residual kind of this request (-1: the case's)
    unsigned cb_ops = 0, cb_copy = 0;     // RK_CALLBACK: the operations and copies to run
    // results
    int status = PFEM_OK;
    pfem_step_stats st{};
    pfem_solve_stats ss{};  // the last linear solve
    double next_dt = 0;
    int64_t t_arrival = 0, t_admit = 0, t_done = 0;
    bool completed = false;
    // engine state
    int slot = -1;
    unsigned ops = 0;  // operations to enqueue at the next tick (OP_* in engine.cpp)
    unsigned copy = 0; // COPY_* bits of the copy operation
    int wait = 0;      // what the readback after those operations decides
    bool final_ = false;
    bool arena_ws = false;
    int iter = 0;
    double res = 0;
    bool verifying = false, fallback = false;
    bool fp64_only = false;  // mixed precision gave up on FP32 for the rest of the request
    int restarts = 0;
    int kit_prev = -1;
    double krr_prev = 0;
    int ref = 0, stall = 0;
    double best = 0;
    bool first_asm = true;
};

class Engine {
public:
    struct Impl;
    Engine(int device, hipStream_t s, int slots, Arena* arena);
    ~Engine();
    Engine(const Engine&) = delete;
    Engine& operator=(const Engine&) = delete;

    // Queues a request; it is admitted (FIFO) when a slot and its workspace are free.
    void enqueue(Req* r);
    // One tick; completed requests are appended to `done`.
    void tick(std::vector<Req*>& done);
    // Ticks until every queued and in-flight request has completed.
    void drain(std::vector<Req*>& done);
    bool idle() const { return waitq_.empty() && inflight_.empty(); }
    int inflight() const { return (int)inflight_.size(); }
    // Grows the slot count (only while idle).
    void ensure_slots(int s);
    // Fails every queued and in-flight request with `code` (after a device error).
    void abort_all(int code, std::vector<Req*>& done);

    void enable_profiling(bool on);
    // captured iteration chunks: {graph launches, captures}
    void graph_stats(long long out[2]) const;
    double prof_ms[2] = {0, 0};
    long long prof_count[2] = {0, 0};
    long long ticks = 0;
    int peak_inflight = 0;
    int max_chunk = 64;

private:
    std::unique_ptr<Impl> p_;
    std::deque<Req*> waitq_;
    std::vector<Req*> inflight_;
};

int64_t now_ns();
// workgroups per case of the row launches (Krylov, residual, refinement); nsl = the
// case's ELL slices (0: none)
int row_blocks(long long n, int nsl = 0);
int pick_lanes(double avg_deg);

}  // namespace pfem

struct pfem_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::unique_ptr<pfem::Engine> engine;  // pfem_batch_step
    std::mutex mu;
};

struct pfem_case {
    // smallest element shape quality 6 sqrt(2) |e| / l_rms^3 of the last mesh (3D; 1 for a regular
    // tet) and the multigrid smoother damping cap it implies (see pfem_case_set_mesh)
    double min_quality = 1.0, amg_omega_cap = 1e30;
    bool amg_all_fp32 = false;  // PFEM_AMG_FP32=1 at set_mesh: 3D mixed hierarchies wholly in FP32
    double setup_s[4] = {0, 0, 0, 0};  // last pfem_case_set_mesh: pattern, ELL, multigrid, total (seconds)
    pfem_ctx* ctx = nullptr;
    pfem_case_desc d{};
    hipStream_t stream = nullptr;
    int dim = 0, N = 0, E = 0, n = 0;
    bool has_mesh = false;
    pfem::Pattern pat;
    // persistent device data
    pfem::DevBuf pos, dir_vals, q_prev;
    pfem::DevBuf partials, counters;
    pfem::DevBuf wstate, wacc;
    pfem::DevBuf wrec;   // WC: [N][8] node records read by the element gathers (WcView::rec)
    pfem::DevBuf werec;  // WC: [E][4] continuity element records (WcView::erec)
    pfem::Workspace ws;
    pfem::AmgHierarchy amg;  // levels == 0: block-Jacobi preconditioning
    int lanes = 8;
    // a node flagged boundary + Dirichlet but not fixed: its position, hence its Dirichlet
    // value, changes at every Picard re-assembly (MomContEquationPSPG.inl:206-214)
    bool moving_dirichlet = false;
    // WC: the CFL minimum dt^2 that the last step (or CFL pass) left for securityCoeff cfl_sc,
    // valid until the state or the positions change (pfem_wc_next_dt then needs no launch)
    bool cfl_valid = false;
    double cfl_dt2 = 0.0, cfl_sc = 0.5;
    // heat (conduction) cases: scalar CSR over the node graph, element records, facets, CG vectors
    struct HeatBufs {
        pfem::DevBuf row_ptr, col, diag_pos, vals, b, erec, facets, fout, fkind, qgp, fterm, nf_ptr, nf_list;
        pfem::DevBuf x, r, z, p, q, dinv, part, scal;
        pfem::DevBuf T, Tbc;  // node temperatures (the state) and Dirichlet temperatures
        uint8_t dbit = 16;    // node flag of the Dirichlet temperatures (PFEM_NODE_DIRICHLET or _T)
        long long nnz = 0;
        int F = 0;
        bool has_q = false;
    } heat;
    // surface-tension facets (pfem_case_set_facets): nodes, elements, node -> (facet, local index)
    struct FstBufs {
        pfem::DevBuf nodes, elem, ptr, ent;
        int F = 0;
    } fst;
    // fractional-step cases: element records, the three graph matrices, right-hand sides, the
    // Picard iterate and the CG vectors
    struct FracBufs {
        pfem::DevBuf erec, mk, ms, lv, dinv_mk, dinv_m, dinv_l, bv, bp, vt, pn, dv, v_it, v_itp, p_itp, pos_save,
            res_part;
        pfem::DevBuf r, z, p, q, part, scal;
    } frac;
    std::atomic<int> busy{0};  // a serving request holds the case
    std::unique_ptr<pfem::Engine> engine;  // direct calls
    ~pfem_case();
};

namespace pfem {
// cases stepped by the PSPG Picard engine (a Boussinesq case adds the heat equation)
inline bool pspg_like(const pfem_case* c) {
    return c->d.kind == PFEM_KIND_PSPG || c->d.kind == PFEM_KIND_BOUSSINESQ;
}
// cases that hold the heat equation
inline bool heat_like(const pfem_case* c) {
    return c->d.kind == PFEM_KIND_HEAT || c->d.kind == PFEM_KIND_BOUSSINESQ;
}
}  // namespace pfem
